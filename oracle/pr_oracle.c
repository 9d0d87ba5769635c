/*
 * pr_oracle.c -- CPU restatement of the reference PageRank (TEST INFRASTRUCTURE).
 *
 * This file is the oracle: only the tests, the smoke check and the CPU
 * baseline leg of the benchmark may load it.  The product path never calls
 * it and has no CPU fallback to it.
 *
 * Graph layout is the reference's adjacency<1> ("in-CSR"): row i lists the
 * in-neighbours j of vertex i (edges j -> i); offsets has n+1 entries
 * (containers/compressed.hpp:67-69, build.hpp:209-265).  deg[j] is the
 * divisor of vertex j's contribution, as passed by the caller
 * (bench/common.hpp:109-120 builds it as the out-degree).
 *
 * Functions and what they follow:
 *   orc_pr_jacobi_f64   page_rank_v1 / page_rank_v10
 *                       (experimental/algorithms/page_rank.hpp:137-174,
 *                        :450-505): contributions recomputed from the
 *                       previous iterate, then r_i = base + d*sum, error =
 *                       sum |r_new - r_old| in double, stop when error < tol
 *                       after a sweep.  Extended with a teleport vector and a
 *                       dangling-mass mode (LEAK = the reference).
 *   orc_pr_gs_f64/_f32  page_rank v11 run on one thread
 *                       (algorithms/page_rank.hpp:89-148): contributions
 *                       initialised once, then updated in place inside the
 *                       sweep (Gauss-Seidel).
 *   orc_pr_gs_par_f32/_f64  the same v11 sweep parallelised with OpenMP
 *                       (schedule(dynamic,64), double reduction), standing in
 *                       for the reference's tbb::parallel_reduce over
 *                       blocked_range(0, N) (page_rank.hpp:124-139); the CPU
 *                       baseline.
 *   orc_pr_jacobi_par_f32   page_rank_v10 with OpenMP (both passes).
 *   orc_pr_bgs_f64      block Gauss-Seidel: rows in a given order cut into
 *                       phases, Jacobi inside a phase, in place between
 *                       phases (between v10 at one phase and serial v11 at
 *                       one row per phase).
 *   orc_wpr_*           weighted PageRank and its adjoint (no reference
 *                       counterpart: the differentiable form of the same
 *                       iteration, see the functions).
 *   orc_pr_delta_f64/_f32  residual PageRank, page_rank_v13 / page_rank_v3
 *                       (experimental/algorithms/page_rank.hpp:566-698).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { ORC_LEAK = 0, ORC_UNIFORM = 1, ORC_TELEPORT = 2 };

/* Jacobi PageRank in double.
 * teleport: NULL => 1/n everywhere. init: NULL => 1/n.
 * dangling: ORC_LEAK (the reference: deg 0 mass is lost), ORC_UNIFORM (mass
 * of deg-0 vertices spread as 1/n), ORC_TELEPORT (spread along teleport).
 * Update: r_i = (1-d)*t_i + d*(sum_{j in row i} r_j/deg_j + D*u_i),
 *   D = sum_{deg_j = 0} r_j.  With LEAK and uniform t this is exactly the
 *   reference's base_score + damping*z (page_rank.hpp:94,133 and
 *   experimental/...:454,492).
 * err_hist (may be NULL) receives the error of each sweep.
 * Returns the number of sweeps performed. */
int orc_pr_jacobi_f64(uint32_t n, const uint32_t* offsets, const uint32_t* indices,
                      const uint32_t* deg, double damping, double tol, uint32_t max_iters,
                      int dangling, const double* teleport, const double* init, double* r,
                      double* err_hist) {
  if (n == 0) return 0;
  double* c = (double*)malloc(sizeof(double) * n);
  double* rn = (double*)malloc(sizeof(double) * n);
  const double un = 1.0 / n;
  for (uint32_t i = 0; i < n; ++i) r[i] = init ? init[i] : un;
  uint32_t it;
  for (it = 0; it < max_iters; ++it) {
    double D = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
      c[i] = r[i] / (double)deg[i]; /* inf for deg 0, never gathered */
      if (deg[i] == 0) D += r[i];
    }
    double err = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
      double z = 0.0;
      for (uint32_t k = offsets[i]; k < offsets[i + 1]; ++k) z += c[indices[k]];
      double t = teleport ? teleport[i] : un;
      double u = dangling == ORC_UNIFORM ? un : (dangling == ORC_TELEPORT ? t : 0.0);
      rn[i] = (1.0 - damping) * t + damping * (z + D * u);
      err += fabs(rn[i] - r[i]);
    }
    memcpy(r, rn, sizeof(double) * n);
    if (err_hist) err_hist[it] = err;
    if (err < tol) {
      ++it;
      break;
    }
  }
  free(c);
  free(rn);
  return (int)it;
}

/* orc_pr_jacobi_f64 with rows split over OpenMP threads.  Each row's sum is
 * still taken sequentially in row order, so the ranks equal the serial
 * function's except for the order of the double reductions of the error and
 * the dangling mass. */
int orc_pr_jacobi_f64_par(uint32_t n, const uint32_t* offsets, const uint32_t* indices,
                          const uint32_t* deg, double damping, double tol, uint32_t max_iters,
                          int dangling, const double* teleport, const double* init, double* r,
                          double* err_hist, int nthreads) {
  if (n == 0) return 0;
  if (nthreads <= 0) nthreads = 1;
  double* c = (double*)malloc(sizeof(double) * n);
  double* rn = (double*)malloc(sizeof(double) * n);
  const double un = 1.0 / n;
  const int64_t N = n;
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (int64_t i = 0; i < N; ++i) r[i] = init ? init[i] : un;
  uint32_t it;
  for (it = 0; it < max_iters; ++it) {
    double D = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : D) num_threads(nthreads)
    for (int64_t i = 0; i < N; ++i) {
      c[i] = r[i] / (double)deg[i];
      if (deg[i] == 0) D += r[i];
    }
    double err = 0.0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : err) num_threads(nthreads)
    for (int64_t i = 0; i < N; ++i) {
      double z = 0.0;
      for (uint32_t k = offsets[i]; k < offsets[i + 1]; ++k) z += c[indices[k]];
      double t = teleport ? teleport[i] : un;
      double u = dangling == ORC_UNIFORM ? un : (dangling == ORC_TELEPORT ? t : 0.0);
      rn[i] = (1.0 - damping) * t + damping * (z + D * u);
      err += fabs(rn[i] - r[i]);
    }
    double* tmp = r;
    memcpy(tmp, rn, sizeof(double) * n);
    if (err_hist) err_hist[it] = err;
    if (err < tol) {
      ++it;
      break;
    }
  }
  free(c);
  free(rn);
  return (int)it;
}

/* Serial v11: c initialised once from r, then in-place Gauss-Seidel sweeps
 * (page_rank.hpp:96-147).  LEAK mode, uniform teleport, as the reference. */
#define ORC_GS_SERIAL(NAME, REAL)                                                              \
  int NAME(uint32_t n, const uint32_t* offsets, const uint32_t* indices, const uint32_t* deg,  \
           REAL damping, REAL tol, uint32_t max_iters, REAL* r, double* err_hist) {            \
    if (n == 0) return 0;                                                                      \
    REAL init_score = (REAL)(1.0 / n);                                                         \
    REAL base_score = (REAL)((1.0 - damping) / n);                                             \
    REAL* c = (REAL*)malloc(sizeof(REAL) * n);                                                 \
    for (uint32_t i = 0; i < n; ++i) r[i] = init_score;                                        \
    for (uint32_t i = 0; i < n; ++i) c[i] = r[i] / (REAL)deg[i];                               \
    uint32_t it;                                                                               \
    for (it = 0; it < max_iters; ++it) {                                                       \
      double err = 0.0;                                                                        \
      for (uint32_t i = 0; i < n; ++i) {                                                       \
        REAL z = 0;                                                                            \
        for (uint32_t k = offsets[i]; k < offsets[i + 1]; ++k) z += c[indices[k]];             \
        REAL old = r[i];                                                                       \
        r[i] = base_score + damping * z;                                                       \
        err += fabs((double)(r[i] - old));                                                     \
        c[i] = r[i] / (REAL)deg[i];                                                            \
      }                                                                                        \
      if (err_hist) err_hist[it] = err;                                                        \
      if (err < tol) {                                                                         \
        ++it;                                                                                  \
        break;                                                                                 \
      }                                                                                        \
    }                                                                                          \
    free(c);                                                                                   \
    return (int)it;                                                                            \
  }

ORC_GS_SERIAL(orc_pr_gs_f64, double)
ORC_GS_SERIAL(orc_pr_gs_f32, float)

/* Block Gauss-Seidel in double: the rows visited in order[0 .. n) are cut
 * into `phases` consecutive phases at cuts[0] = 0 < ... <= cuts[phases] = n
 * (positions in `order`).  Inside a phase every row gathers the
 * contributions as they stood when the phase began (Jacobi); a phase's new
 * ranks and contributions replace the old ones before the next phase
 * gathers (Gauss-Seidel between phases).  phases = 1 is the Jacobi sweep of
 * orc_pr_jacobi_f64; one row per phase is the serial in-place sweep of
 * page_rank v11 (page_rank.hpp:124-139) in that row order.  Update, stop rule
 * and sweep count as orc_pr_jacobi_f64 with LEAK dangling (the reference's:
 * no dangling mass is redistributed), error = sum over all phases of
 * |r_new - r_old|. */
int orc_pr_bgs_f64(uint32_t n, const uint32_t* offsets, const uint32_t* indices, const uint32_t* deg,
                   double damping, double tol, uint32_t max_iters, const uint32_t* order, uint32_t phases,
                   const uint32_t* cuts, const double* teleport, const double* init, double* r,
                   double* err_hist, int nthreads) {
  if (n == 0) return 0;
  if (nthreads <= 0) nthreads = 1;
  double* c = (double*)malloc(sizeof(double) * n);
  double* rn = (double*)malloc(sizeof(double) * n);
  const double un = 1.0 / n;
  for (uint32_t i = 0; i < n; ++i) r[i] = init ? init[i] : un;
  for (uint32_t i = 0; i < n; ++i) c[i] = r[i] / (double)deg[i]; /* inf for deg 0, never gathered */
  uint32_t it;
  for (it = 0; it < max_iters; ++it) {
    double err = 0.0;
    for (uint32_t q = 0; q < phases; ++q) {
      /* the rows of one phase are independent: threads change only the
         order of the error sum, not the ranks */
      const int64_t lo = cuts[q], hi = cuts[q + 1];
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads) if (hi - lo > 4096)
      for (int64_t pos = lo; pos < hi; ++pos) {
        const uint32_t i = order[pos];
        double z = 0.0;
        for (uint32_t k = offsets[i]; k < offsets[i + 1]; ++k) z += c[indices[k]];
        const double t = teleport ? teleport[i] : un;
        rn[i] = (1.0 - damping) * t + damping * z;
      }
      double e = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : e) num_threads(nthreads) if (hi - lo > 4096)
      for (int64_t pos = lo; pos < hi; ++pos) {
        const uint32_t i = order[pos];
        e += fabs(rn[i] - r[i]);
        r[i] = rn[i];
        c[i] = r[i] / (double)deg[i];
      }
      err += e;
    }
    if (err_hist) err_hist[it] = err;
    if (err < tol) {
      ++it;
      break;
    }
  }
  free(c);
  free(rn);
  return (int)it;
}

/* Parallel v11 (the CPU baseline): the same in-place sweep with the vertex
 * range split over OpenMP threads, as tbb::parallel_reduce splits
 * blocked_range(0, N) (page_rank.hpp:124-139).  The residual reduction is in
 * double, as the reference's 0.0 identity.  Reads of c[] by one thread may see
 * another thread's updates of the same sweep, exactly as in the reference. */
#define ORC_GS_PAR(NAME, REAL)                                                                   \
  int NAME(uint32_t n, const uint32_t* offsets, const uint32_t* indices, const uint32_t* deg,    \
           REAL damping, REAL tol, uint32_t max_iters, int nthreads, REAL* r, double* err_hist) { \
    if (n == 0) return 0;                                                                        \
    if (nthreads <= 0) nthreads = 1;                                                             \
    REAL init_score = (REAL)(1.0 / n);                                                           \
    REAL base_score = (REAL)((1.0 - damping) / n);                                               \
    REAL* c = (REAL*)malloc(sizeof(REAL) * n);                                                   \
    int64_t N = n;                                                                               \
    _Pragma("omp parallel for schedule(static) num_threads(nthreads)")                          \
    for (int64_t i = 0; i < N; ++i) r[i] = init_score;                                           \
    _Pragma("omp parallel for schedule(static) num_threads(nthreads)")                          \
    for (int64_t i = 0; i < N; ++i) c[i] = r[i] / (REAL)deg[i];                                  \
    uint32_t it;                                                                                 \
    for (it = 0; it < max_iters; ++it) {                                                         \
      double err = 0.0;                                                                          \
      _Pragma("omp parallel for schedule(dynamic, 64) reduction(+ : err) num_threads(nthreads)") \
      for (int64_t i = 0; i < N; ++i) {                                                          \
        REAL z = 0;                                                                              \
        for (uint32_t k = offsets[i]; k < offsets[i + 1]; ++k) z += c[indices[k]];               \
        REAL old = r[i];                                                                         \
        REAL nr = base_score + damping * z;                                                      \
        r[i] = nr;                                                                               \
        err += fabs((double)(nr - old));                                                         \
        c[i] = nr / (REAL)deg[i];                                                                \
      }                                                                                          \
      if (err_hist) err_hist[it] = err;                                                          \
      if (err < tol) {                                                                           \
        ++it;                                                                                    \
        break;                                                                                   \
      }                                                                                          \
    }                                                                                            \
    free(c);                                                                                     \
    return (int)it;                                                                              \
  }

ORC_GS_PAR(orc_pr_gs_par_f32, float)
ORC_GS_PAR(orc_pr_gs_par_f64, double)

/* Parallel v10 Jacobi (experimental/algorithms/page_rank.hpp:450-505). */
int orc_pr_jacobi_par_f32(uint32_t n, const uint32_t* offsets, const uint32_t* indices,
                          const uint32_t* deg, float damping, float tol, uint32_t max_iters,
                          int nthreads, float* r, double* err_hist) {
  if (n == 0) return 0;
  if (nthreads <= 0) nthreads = 1;
  float init_score = (float)(1.0 / n);
  float base_score = (float)((1.0 - damping) / n);
  float* c = (float*)malloc(sizeof(float) * n);
  int64_t N = n;
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (int64_t i = 0; i < N; ++i) r[i] = init_score;
  uint32_t it;
  for (it = 0; it < max_iters; ++it) {
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int64_t i = 0; i < N; ++i) c[i] = r[i] / (float)deg[i];
    double err = 0.0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : err) num_threads(nthreads)
    for (int64_t i = 0; i < N; ++i) {
      float z = 0;
      for (uint32_t k = offsets[i]; k < offsets[i + 1]; ++k) z += c[indices[k]];
      float old = r[i];
      r[i] = base_score + damping * z;
      err += fabs((double)(r[i] - old));
    }
    if (err_hist) err_hist[it] = err;
    if (err < tol) {
      ++it;
      break;
    }
  }
  free(c);
  return (int)it;
}

/* ---------------------------------------------------------------------------
 * Weighted PageRank and its adjoint (double).
 *
 * Edge k of row v (in-CSR) is u -> v with raw weight w[k] > 0.  s_u = sum of
 * w over u's out-edges; P_uv = w_uv / s_u.  Fixed point
 *   r_v = (1-d)*t_v + d*(sum_{u->v} P_uv r_u + D*q_v),  D = sum_{s_u = 0} r_u,
 * q = 0 (LEAK), 1/n (UNIFORM) or t (TELEPORT).  With w = 1 this is the plain
 * Jacobi iteration above with deg = out-degree.
 *
 * Adjoint of L(r*) with g = dL/dr at the fixed point:
 *   lambda = d*(P lambda + 1_dangling * (q . lambda)) + g    (solved by
 *   fixed-point iteration, contraction d), then
 *   dL/dt_v  = (1-d)*lambda_v + d*D_t,   D_t = (q == t) ? D*lambda_v : 0
 *   dL/dw_uv = d * r_u * (lambda_v - (P lambda)_u) / s_u
 * (derivation: dF/dP_uv = d*lambda_v*r_u and dP_uv'/dw_uv = (delta - P_uv')/s_u).
 * ------------------------------------------------------------------------- */
static void wpr_out_sums(uint32_t n, const uint32_t* offsets, const uint32_t* indices,
                         const double* w, double* s) {
  memset(s, 0, sizeof(double) * n);
  for (uint32_t v = 0; v < n; ++v)
    for (uint32_t k = offsets[v]; k < offsets[v + 1]; ++k) s[indices[k]] += w[k];
}

int orc_wpr_forward_f64(uint32_t n, const uint32_t* offsets, const uint32_t* indices,
                        const double* w, const double* t, double damping, double tol,
                        uint32_t max_iters, int dangling, double* r, double* err_hist) {
  if (n == 0) return 0;
  double* s = (double*)malloc(sizeof(double) * n);
  double* c = (double*)malloc(sizeof(double) * n);
  double* rn = (double*)malloc(sizeof(double) * n);
  wpr_out_sums(n, offsets, indices, w, s);
  const double un = 1.0 / n;
  for (uint32_t i = 0; i < n; ++i) r[i] = un;
  uint32_t it;
  for (it = 0; it < max_iters; ++it) {
    double D = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
      c[i] = s[i] > 0 ? r[i] / s[i] : 0.0;
      if (!(s[i] > 0)) D += r[i];
    }
    double err = 0.0;
    for (uint32_t v = 0; v < n; ++v) {
      double z = 0.0;
      for (uint32_t k = offsets[v]; k < offsets[v + 1]; ++k) z += w[k] * c[indices[k]];
      double q = dangling == ORC_UNIFORM ? un : (dangling == ORC_TELEPORT ? t[v] : 0.0);
      rn[v] = (1.0 - damping) * t[v] + damping * (z + D * q);
      err += fabs(rn[v] - r[v]);
    }
    memcpy(r, rn, sizeof(double) * n);
    if (err_hist) err_hist[it] = err;
    if (err < tol) {
      ++it;
      break;
    }
  }
  free(s);
  free(c);
  free(rn);
  return (int)it;
}

/* lambda solve + gradients; r is the forward fixed point, g = dL/dr.
 * Writes lambda (n), dLdw (m, in-CSR edge order), dLdt (n).
 * Returns the number of adjoint sweeps. */
int orc_wpr_adjoint_f64(uint32_t n, const uint32_t* offsets, const uint32_t* indices,
                        const double* w, const double* t, const double* r, const double* g,
                        double damping, double tol, uint32_t max_iters, int dangling,
                        double* lambda, double* dLdw, double* dLdt) {
  if (n == 0) return 0;
  double* s = (double*)malloc(sizeof(double) * n);
  double* pl = (double*)malloc(sizeof(double) * n);
  double* ln = (double*)malloc(sizeof(double) * n);
  wpr_out_sums(n, offsets, indices, w, s);
  const double un = 1.0 / n;
  for (uint32_t i = 0; i < n; ++i) lambda[i] = g[i];
  uint32_t it;
  for (it = 0; it < max_iters; ++it) {
    /* pl_u = sum_{u->v} P_uv lambda_v */
    memset(pl, 0, sizeof(double) * n);
    for (uint32_t v = 0; v < n; ++v)
      for (uint32_t k = offsets[v]; k < offsets[v + 1]; ++k) {
        uint32_t u = indices[k];
        pl[u] += w[k] / s[u] * lambda[v];
      }
    double ql = 0.0;
    if (dangling == ORC_UNIFORM)
      for (uint32_t v = 0; v < n; ++v) ql += un * lambda[v];
    else if (dangling == ORC_TELEPORT)
      for (uint32_t v = 0; v < n; ++v) ql += t[v] * lambda[v];
    double err = 0.0;
    for (uint32_t u = 0; u < n; ++u) {
      double x = s[u] > 0 ? pl[u] : ql;
      ln[u] = damping * x + g[u];
      err += fabs(ln[u] - lambda[u]);
    }
    memcpy(lambda, ln, sizeof(double) * n);
    if (err < tol) {
      ++it;
      break;
    }
  }
  /* final P lambda with the converged lambda */
  memset(pl, 0, sizeof(double) * n);
  for (uint32_t v = 0; v < n; ++v)
    for (uint32_t k = offsets[v]; k < offsets[v + 1]; ++k) {
      uint32_t u = indices[k];
      pl[u] += w[k] / s[u] * lambda[v];
    }
  for (uint32_t v = 0; v < n; ++v)
    for (uint32_t k = offsets[v]; k < offsets[v + 1]; ++k) {
      uint32_t u = indices[k];
      dLdw[k] = damping * r[u] * (lambda[v] - pl[u]) / s[u];
    }
  double D = 0.0;
  for (uint32_t u = 0; u < n; ++u)
    if (!(s[u] > 0)) D += r[u];
  for (uint32_t v = 0; v < n; ++v)
    dLdt[v] = (1.0 - damping) * lambda[v] + (dangling == ORC_TELEPORT ? damping * D * lambda[v] : 0.0);
  free(s);
  free(pl);
  free(ln);
  return (int)it;
}

/* Weighted forward and adjoint with rows split over OpenMP threads.  The
 * adjoint gathers over an out-CSR built here (row u lists the in-CSR
 * positions of u's out-edges in ascending order), so each (P lambda)_u is
 * summed sequentially and the result does not depend on the thread count. */
int orc_wpr_forward_f64_par(uint32_t n, const uint32_t* offsets, const uint32_t* indices,
                            const double* w, const double* t, double damping, double tol,
                            uint32_t max_iters, int dangling, const double* init, double* r,
                            int nthreads) {
  if (n == 0) return 0;
  if (nthreads <= 0) nthreads = 1;
  double* s = (double*)malloc(sizeof(double) * n);
  double* c = (double*)malloc(sizeof(double) * n);
  double* rn = (double*)malloc(sizeof(double) * n);
  wpr_out_sums(n, offsets, indices, w, s);
  const double un = 1.0 / n;
  const int64_t N = n;
  for (uint32_t i = 0; i < n; ++i) r[i] = init ? init[i] : un;
  uint32_t it;
  for (it = 0; it < max_iters; ++it) {
    double D = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : D) num_threads(nthreads)
    for (int64_t i = 0; i < N; ++i) {
      c[i] = s[i] > 0 ? r[i] / s[i] : 0.0;
      if (!(s[i] > 0)) D += r[i];
    }
    double err = 0.0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : err) num_threads(nthreads)
    for (int64_t v = 0; v < N; ++v) {
      double z = 0.0;
      for (uint32_t k = offsets[v]; k < offsets[v + 1]; ++k) z += w[k] * c[indices[k]];
      double q = dangling == ORC_UNIFORM ? un : (dangling == ORC_TELEPORT ? t[v] : 0.0);
      rn[v] = (1.0 - damping) * t[v] + damping * (z + D * q);
      err += fabs(rn[v] - r[v]);
    }
    memcpy(r, rn, sizeof(double) * n);
    if (err < tol) {
      ++it;
      break;
    }
  }
  free(s);
  free(c);
  free(rn);
  return (int)it;
}

int orc_wpr_adjoint_f64_par(uint32_t n, const uint32_t* offsets, const uint32_t* indices,
                            const double* w, const double* t, const double* r, const double* g,
                            double damping, double tol, uint32_t max_iters, int dangling,
                            double* lambda, double* dLdw, double* dLdt, int nthreads) {
  if (n == 0) return 0;
  if (nthreads <= 0) nthreads = 1;
  const uint32_t m = offsets[n];
  double* s = (double*)malloc(sizeof(double) * n);
  double* pl = (double*)malloc(sizeof(double) * n);
  double* ln = (double*)malloc(sizeof(double) * n);
  uint32_t* ooff = (uint32_t*)calloc((size_t)n + 1, sizeof(uint32_t));
  uint32_t* opos = (uint32_t*)malloc(sizeof(uint32_t) * (m ? m : 1));
  uint32_t* orow = (uint32_t*)malloc(sizeof(uint32_t) * (m ? m : 1));
  wpr_out_sums(n, offsets, indices, w, s);
  for (uint32_t k = 0; k < m; ++k) ooff[indices[k] + 1]++;
  for (uint32_t u = 0; u < n; ++u) ooff[u + 1] += ooff[u];
  {
    uint32_t* cur = (uint32_t*)malloc(sizeof(uint32_t) * n);
    memcpy(cur, ooff, sizeof(uint32_t) * n);
    for (uint32_t v = 0; v < n; ++v)
      for (uint32_t k = offsets[v]; k < offsets[v + 1]; ++k) {
        uint32_t p = cur[indices[k]]++;
        opos[p] = k;
        orow[p] = v;
      }
    free(cur);
  }
  const double un = 1.0 / n;
  const int64_t N = n;
  for (uint32_t i = 0; i < n; ++i) lambda[i] = g[i];
  uint32_t it;
  for (it = 0; it < max_iters; ++it) {
    double ql = 0.0;
    if (dangling == ORC_UNIFORM)
      for (uint32_t v = 0; v < n; ++v) ql += un * lambda[v];
    else if (dangling == ORC_TELEPORT)
      for (uint32_t v = 0; v < n; ++v) ql += t[v] * lambda[v];
    double err = 0.0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : err) num_threads(nthreads)
    for (int64_t u = 0; u < N; ++u) {
      double z = 0.0;
      for (uint32_t p = ooff[u]; p < ooff[u + 1]; ++p) z += w[opos[p]] / s[u] * lambda[orow[p]];
      pl[u] = z;
      double x = s[u] > 0 ? z : ql;
      ln[u] = damping * x + g[u];
      err += fabs(ln[u] - lambda[u]);
    }
    memcpy(lambda, ln, sizeof(double) * n);
    if (err < tol) {
      ++it;
      break;
    }
  }
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
  for (int64_t u = 0; u < N; ++u) {
    double z = 0.0;
    for (uint32_t p = ooff[u]; p < ooff[u + 1]; ++p) z += w[opos[p]] / s[u] * lambda[orow[p]];
    pl[u] = z;
  }
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
  for (int64_t v = 0; v < N; ++v)
    for (uint32_t k = offsets[v]; k < offsets[v + 1]; ++k) {
      uint32_t u = indices[k];
      dLdw[k] = damping * r[u] * (lambda[v] - pl[u]) / s[u];
    }
  double D = 0.0;
  for (uint32_t u = 0; u < n; ++u)
    if (!(s[u] > 0)) D += r[u];
  for (uint32_t v = 0; v < n; ++v)
    dLdt[v] = (1.0 - damping) * lambda[v] + (dangling == ORC_TELEPORT ? damping * D * lambda[v] : 0.0);
  free(s);
  free(pl);
  free(ln);
  free(ooff);
  free(opos);
  free(orow);
  return (int)it;
}

/* Residual ("delta") PageRank: page_rank_v13
 * (experimental/algorithms/page_rank.hpp:632-698) with keep = 0, and the
 * serial page_rank_v3 (:566-630) with keep = 1.  Ranks start at 0 and every
 * vertex's residual at 1 - d.  Each iteration:
 *   phase 1: a vertex whose residual exceeds `threshold` adds it to its rank
 *            and, if its degree is non-zero, sends delta = res * d / deg
 *            (its residual becomes 0; a vertex that does not push sends 0);
 *   phase 2: sum_i = sum over row i of the positive deltas; residual_i =
 *            sum_i if sum_i > 0, else 0 (v13) or unchanged (v3: a residual
 *            at or below the threshold waits for later deltas, but is
 *            replaced, not added to, when one arrives);
 * the loop stops after the first iteration whose phase 1 pushed nothing, or
 * after max_iters iterations.  Ranks are not normalised (they sum to about n
 * times the leaked-mass fraction).  Real = float (f32) or double (f64), with
 * the reference's operation order; row sums in row order.  pushes[it] (when
 * given, max_iters entries) receives the number of pushing vertices of
 * non-zero degree in iteration it + 1 (the reference's `changed` is this
 * count > 0).
 * Returns the iterations performed (phase 1 + phase 2 pairs). */
#define ORC_DELTA_BODY(REAL)                                                                       \
  if (n == 0 || max_iters == 0) {                                                                  \
    for (uint32_t i = 0; i < n; ++i) rank[i] = 0;                                                  \
    return 0;                                                                                      \
  }                                                                                                \
  REAL* delta = (REAL*)malloc(sizeof(REAL) * n);                                                   \
  REAL* res = (REAL*)malloc(sizeof(REAL) * n);                                                     \
  if (!delta || !res) {                                                                            \
    free(delta);                                                                                   \
    free(res);                                                                                     \
    return -1;                                                                                     \
  }                                                                                                \
  const REAL d = (REAL)damping, thr = (REAL)threshold;                                             \
  for (uint32_t i = 0; i < n; ++i) {                                                               \
    rank[i] = 0;                                                                                   \
    delta[i] = 0;                                                                                  \
    res[i] = (REAL)(1.0 - (double)d);                                                              \
  }                                                                                                \
  uint32_t it = 0;                                                                                 \
  for (; it < max_iters;) {                                                                        \
    int sends = 0;                                                                                 \
    for (uint32_t s = 0; s < n; ++s) {                                                             \
      delta[s] = 0;                                                                                \
      if (res[s] > thr) {                                                                          \
        const REAL old = res[s];                                                                   \
        rank[s] += old;                                                                            \
        res[s] = 0;                                                                                \
        if (deg[s] > 0) {                                                                          \
          delta[s] = old * d / (REAL)deg[s];                                                       \
          ++sends;                                                                                 \
        }                                                                                          \
      }                                                                                            \
    }                                                                                              \
    for (uint32_t i = 0; i < n; ++i) {                                                             \
      REAL sum = 0;                                                                                \
      for (uint32_t k = offsets[i]; k < offsets[i + 1]; ++k)                                       \
        if (delta[indices[k]] > 0) sum += delta[indices[k]];                                       \
      if (sum > 0)                                                                                 \
        res[i] = sum;                                                                              \
      else if (!keep)                                                                              \
        res[i] = 0;                                                                                \
    }                                                                                              \
    ++it;                                                                                          \
    if (pushes) pushes[it - 1] = sends;                                                            \
    if (!sends) break;                                                                             \
  }                                                                                                \
  free(delta);                                                                                     \
  free(res);                                                                                       \
  return (int)it;

int orc_pr_delta_f64(uint32_t n, const uint32_t* offsets, const uint32_t* indices, const uint32_t* deg,
                     double damping, double threshold, uint32_t max_iters, int keep, double* rank, int* pushes) {
  ORC_DELTA_BODY(double)
}

int orc_pr_delta_f32(uint32_t n, const uint32_t* offsets, const uint32_t* indices, const uint32_t* deg,
                     double damping, double threshold, uint32_t max_iters, int keep, float* rank, int* pushes) {
  ORC_DELTA_BODY(float)
}
