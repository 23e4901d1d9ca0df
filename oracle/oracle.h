/*
 * oracle.h -- CPU restatement of the reference's agent-step semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or the timed CPU baseline.  The shipped engine (jaxabm_amd) never links it.
 *
 * Restates:
 *   - JAX's threefry2x32 PRNG construction (key derivation, split, random
 *     bits, uniform, randint, normal, permutation) in both stream modes,
 *     as called from jaxabm/model.py:46,129,156,164,183 and
 *     jaxabm/agent.py:115,156 (JAX itself is an un-vendored dependency,
 *     requirements.txt:9-10, jax>=0.4.1 unpinned).
 *   - The reference step schedule of JaxModel.initialize/step
 *     (jaxabm/model.py:118-216) applied to five model programs written
 *     against the reference API: wealth exchange, Schelling, network SIR,
 *     Deffuant (forward and adjoint), and a generic trader model (the
 *     per-agent code of the generic-model workload).
 *   - jax.random's gamma / loggamma / beta samplers (Marsaglia-Tsang).
 *   - Record-time reductions: mean, Gini (examples/models/
 *     advanced_economic_model.py:1798-1811), S/I/R counts, soft histogram.
 *
 * Stream modes: ORC_MODE_ORIGINAL = jax_threefry_partitionable=False
 * (JAX < 0.5 default); ORC_MODE_PARTITIONABLE = True (JAX >= 0.5 default).
 */
#ifndef JAXABM_ORACLE_H
#define JAXABM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MODE_ORIGINAL 0
#define ORC_MODE_PARTITIONABLE 1

/* ---- PRNG ---- */
void orc_threefry2x32(uint32_t k0, uint32_t k1, uint32_t x0, uint32_t x1,
                      uint32_t out[2]);
void orc_prng_key(int64_t seed, uint32_t key[2]);
void orc_split(const uint32_t key[2], int64_t num, uint32_t *out, int mode);
void orc_random_bits32(const uint32_t key[2], int64_t n, uint32_t *out,
                       int mode);
void orc_uniform_f32(const uint32_t key[2], int64_t n, float lo, float hi,
                     float *out, int mode);
void orc_randint_i32(const uint32_t key[2], int64_t n, int32_t lo, int32_t hi,
                     int32_t *out, int mode);
/* Elementwise float32 exp (fn 0), log (1), log1p (2) and pow(x, y) (3), and
 * float64 exp (0) and log (1), from the shared abm_math.h (y may be NULL
 * unless fn is pow). */
void orc_math_f32(int fn, const float *x, const float *y, float *out, int64_t n);
void orc_math_f64(int fn, const double *x, double *out, int64_t n);
float orc_erfinv_f32(float x);
void orc_normal_f32(const uint32_t key[2], int64_t n, float *out, int mode);
/* gamma(key, alpha[j], (count,)) (log_space: loggamma) and
 * beta(key, a[j], b[j], (count,)), float32. */
void orc_gamma_f32(const uint32_t key[2], int64_t count, const float *alpha, int log_space,
                   float *out, int mode);
void orc_beta_f32(const uint32_t key[2], int64_t count, const float *a, const float *b,
                  float *out, int mode);
int32_t orc_permutation_rounds(int64_t n);
void orc_permutation(const uint32_t key[2], int64_t n, int32_t *out, int mode);

/* ---- network ---- */
int64_t orc_ba_num_edges(int64_t n, int32_t m);
int64_t orc_ba_edges(int64_t n, int32_t m, const uint32_t key[2],
                     int32_t *src, int32_t *dst);
void orc_csr_undirected(int64_t n, int64_t n_edges, const int32_t *src,
                        const int32_t *dst, int32_t *row_ptr, int32_t *col);

/* ---- reductions ---- */
float orc_gini_i32(const int32_t *v, int64_t n);
float orc_mean_i32(const int32_t *v, int64_t n);

/* ---- programs ---- */
void orc_sir_init(int64_t n, int32_t n_seed_infected, int64_t seed, int mode,
                  uint8_t *state, uint32_t rng[2]);
void orc_sir_run(int64_t n, const int32_t *row_ptr, const int32_t *col,
                 uint8_t *state, uint32_t rng[2], float beta, float gamma,
                 int32_t steps, int mode, int32_t *counts);
void orc_sir_run_threads(int64_t n, const int32_t *row_ptr, const int32_t *col,
                         uint8_t *state, uint32_t rng[2], float beta, float gamma,
                         int32_t steps, int mode, int32_t *counts, int threads);

void orc_trader_init(int64_t n, int64_t seed, int mode, float *x, int32_t *w,
                     uint32_t rng[2]);
void orc_trader_run(int64_t n, float *x, int32_t *w, uint32_t rng[2], int32_t steps,
                    int mode);
void orc_wealth_init(int64_t n, int64_t seed, int mode, float *x, float *y,
                     int32_t *wealth, uint32_t rng[2]);
void orc_wealth_run_threads(int64_t n, float *x, float *y, int32_t *wealth, uint32_t rng[2],
                            int32_t steps, int mode, float *mean_out, float *gini_out,
                            int threads);
void orc_wealth_run(int64_t n, float *x, float *y, int32_t *wealth,
                    uint32_t rng[2], int32_t steps, int mode, float *mean_out,
                    float *gini_out);

int64_t orc_schelling_init(int32_t height, int32_t width, int64_t seed,
                           int mode, int8_t *cells, int32_t *agent_pos,
                           int8_t *agent_type, uint32_t rng[2]);
void orc_schelling_run(int32_t height, int32_t width, int8_t *cells,
                       int32_t *agent_pos, const int8_t *agent_type,
                       int64_t n_agents, uint32_t rng[2], int32_t steps,
                       int mode, float *happy_frac_out, int32_t *moves_out);
/* The same run with each step's agent phase on `threads` OpenMP threads
 * (same results; a CPU baseline on all cores). */
void orc_schelling_run_threads(int32_t height, int32_t width, int8_t *cells,
                               int32_t *agent_pos, const int8_t *agent_type,
                               int64_t n_agents, uint32_t rng[2], int32_t steps,
                               int mode, float *happy_frac_out, int32_t *moves_out,
                               int threads);

void orc_deffuant_init(int64_t n, int32_t batch, int64_t seed, int mode,
                       float *x, uint32_t rng[2]);
void orc_deffuant_forward(int64_t n, int32_t batch, const float *eps,
                          const float *mu, float *x, uint32_t rng[2],
                          int32_t steps, int mode);
void orc_deffuant_stats(int64_t n, int32_t batch, const float *x,
                        double *stats);
double orc_deffuant_loss(const double *stats, const double *target);
void orc_deffuant_grad(int64_t n, int32_t batch, const double *eps,
                       const double *mu, const float *x0,
                       const uint32_t rng0[2], int32_t steps, int mode,
                       const double *target, double *loss_out,
                       double *grad_eps, double *grad_mu);
/* The same outputs, bit for bit, with the instances over `threads` OpenMP
 * threads (a CPU baseline on all cores). */
void orc_deffuant_grad_threads(int64_t n, int32_t batch, const double *eps,
                               const double *mu, const float *x0,
                               const uint32_t rng0[2], int32_t steps, int mode,
                               const double *target, double *loss_out,
                               double *grad_eps, double *grad_mu, int32_t threads);

#define ORC_HIST_BINS 32
#define ORC_STATS_LEN (2 + ORC_HIST_BINS)

#ifdef __cplusplus
}
#endif
#endif
