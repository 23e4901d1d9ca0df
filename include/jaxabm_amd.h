/*
 * jaxabm_amd.h -- C-ABI of the MI355X agent-step engine (libjaxabm_hip.so).
 *
 * Plain C: integer status returns (0 = ok), a thread-local message from
 * abm_last_error(), caller-owned device buffers passed as raw pointers with
 * explicit sizes, and an explicit stream (hipStream_t passed as void*).
 * The engine never frees or reallocates caller memory.
 *
 * The reference (a11to1n3/JaxABM) is pure Python over JAX; its "boundary" for
 * this path is the Python protocol listed beside each entry point, whose JAX
 * primitives these functions replace.  Python binds them with ctypes
 * (jaxabm_amd/_lib.py).
 *
 * RNG `mode`: 0 = JAX threefry stream with jax_threefry_partitionable=False
 * (JAX < 0.5 default), 1 = partitionable stream (JAX >= 0.5 default).
 * Keys are uint32[2] in device memory unless the name says host.
 */
#ifndef JAXABM_AMD_H
#define JAXABM_AMD_H
#include <stddef.h>
#include <stdint.h>

#define JAXABM_AMD_ABI_VERSION 1

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status ---- */
/* Message for the last non-zero status on this thread. */
const char *abm_last_error(void);
int abm_version(void);

/* ---- PRNG: replaces jax.random.* at jaxabm/model.py:46,129,156,164,183,
 *      jaxabm/agent.py:115,156, jaxabm/agentpy.py:509-513 ---- */
/* Raw threefry2x32-20 block over n counter pairs (known-answer tests). */
int abm_threefry2x32(const uint32_t *key, const uint32_t *x0, const uint32_t *x1,
                     uint32_t *y0, uint32_t *y1, int64_t n, void *stream);
/* jax.random.split(key, num) -> out[num][2]  (jaxabm/model.py:129,156). */
int abm_split(const uint32_t *key, int64_t num, uint32_t *out, int mode, void *stream);
/* jax.random.bits(key, (n,), uint32). */
int abm_random_bits32(const uint32_t *key, int64_t n, uint32_t *out, int mode, void *stream);
/* jax.random.uniform(key, (n,), float32, lo, hi). */
int abm_uniform_f32(const uint32_t *key, int64_t n, float lo, float hi, float *out,
                    int mode, void *stream);
/* jax.random.normal(key, (n,), float32). */
int abm_normal_f32(const uint32_t *key, int64_t n, float *out, int mode, void *stream);
/* jax.random.bernoulli(key, p, (n,)) as uint8 0/1. */
int abm_bernoulli_u8(const uint32_t *key, int64_t n, float p, uint8_t *out, int mode,
                     void *stream);
/* jax.random.randint(key, (n,), lo, hi, int32). */
int abm_randint_i32(const uint32_t *key, int64_t n, int32_t lo, int32_t hi, int32_t *out,
                    int mode, void *stream);
/* vmap(split(., num))(keys[n][2]) -> out[n][num][2]  (jaxabm/agent.py:156 + vmap). */
int abm_split_each(const uint32_t *keys, int64_t n, int64_t num, uint32_t *out, int mode,
                   void *stream);
/* vmap of a per-key draw of `per` elements over keys[n][2] -> out[n][per].
 * dist: 0 bits (uint32), 1 uniform(lo, hi) (float), 2 normal (float),
 * 3 bernoulli(p = lo) (uint8). */
int abm_draw_each(const uint32_t *keys, int64_t n, int64_t per, int dist, float lo,
                  float hi, void *out, int mode, void *stream);
/* vmap of randint(key, (per,), lo, hi) over keys[n][2] -> out[n][per]. */
int abm_randint_each(const uint32_t *keys, int64_t n, int64_t per, int32_t lo, int32_t hi,
                     int32_t *out, int mode, void *stream);
/* jax.random.permutation(key, n) (jaxabm/analysis.py:87); workspace from
 * abm_permutation_workspace(n). */
size_t abm_permutation_workspace(int64_t n);
int abm_permutation_rounds(int64_t n);
int abm_permutation(const uint32_t *key, int64_t n, int32_t *out, void *ws, size_t ws_bytes,
                    int mode, void *stream);
/* Stable ascending sort of (keys, vals) in place; workspace as permutation. */
int abm_sort_pairs_u32(uint32_t *keys, int32_t *vals, int64_t n, void *ws, size_t ws_bytes,
                       void *stream);
/* Host key arithmetic: split of a host key into host out[num][2]. */
int abm_host_split(const uint32_t key[2], int64_t num, uint32_t *out, int mode);
/* Host key schedule of JaxModel.step (jaxabm/model.py:156,164,183) for
 * `steps` steps, advancing rng[2] in place.  Row per step: the collection
 * keys ck_0..ck_{C-1}, then the env-phase key uk (or split(uk) when
 * split_env != 0). */
int abm_step_keys(uint32_t rng[2], int32_t steps, int32_t n_collections,
                  int32_t has_env_phase, int32_t split_env, int mode, uint32_t *out);

/* ---- Network construction: replaces Network.add_edge/get_neighbors,
 *      jaxabm/agentpy.py:559-615 (host memory) ---- */
int64_t abm_ba_num_edges(int64_t n, int32_t m);
/* Barabasi-Albert edge list; returns the number of edges or -1. */
int64_t abm_ba_edges(int64_t n, int32_t m, const uint32_t key[2], int32_t *src,
                     int32_t *dst);
/* Undirected CSR (both directions, rows ascending). */
int abm_csr_undirected(int64_t n, int64_t n_edges, const int32_t *src, const int32_t *dst,
                       int32_t *row_ptr, int32_t *col);
/* Rows with degree > threshold (ascending); out may be NULL to count. */
int64_t abm_heavy_rows(int64_t n, const int32_t *row_ptr, int32_t threshold, int32_t *out);

/* ---- network SIR program: the env-phase callback of JaxModel.step
 *      (jaxabm/model.py:182-190) + metrics_fn counts (model.py:194-200) ---- */
typedef struct abm_sir_graph {
  int64_t n;              /* agents */
  int64_t nnz;            /* CSR entries (2 x undirected edges) */
  const int32_t *row_ptr; /* device [n+1] */
  const int32_t *col;     /* device [nnz] */
  const int32_t *heavy;   /* device [n_heavy]: rows with degree > abm_sir_heavy_threshold() */
  int64_t n_heavy;
} abm_sir_graph;
int abm_sir_heavy_threshold(void);
/* state[n] = 0; state[perm[i]] = 1 for i < n0. */
int abm_sir_seed(int64_t n, const int32_t *perm, int64_t n0, uint8_t *state, void *stream);
/* Runs `steps` synchronous steps.  keys: device [steps][4] = (k_inf, k_rec)
 * per step (abm_step_keys with split_env).  counts: device [steps][3] S/I/R
 * after each step.  scratch: device [n]. */
int abm_sir_run(const abm_sir_graph *g, uint8_t *state, uint8_t *scratch, const uint32_t *keys,
                int32_t steps, float beta, float gamma, int mode, int32_t *counts, void *stream);
/* One step in -> out (in != out); counts3 must be zeroed by the caller
 * (it accumulates S, I, R). */
int abm_sir_step(const abm_sir_graph *g, const uint8_t *in, uint8_t *out, const uint32_t *keys4,
                 float beta, float gamma, int mode, int32_t *counts3, void *stream);
/* ceil(p * 2^23): integer form of `uniform < p`. */
uint32_t abm_unit_threshold(float p);
int abm_fill_i32(int32_t *p, int64_t n, int32_t v, void *stream);

/* ---- step-loop capture (replaces the Python loop of JaxModel.run,
 *      jaxabm/model.py:240-241) ---- */
typedef struct abm_graph abm_graph;
int abm_graph_begin(void *stream);
int abm_graph_end(void *stream, abm_graph **out);
int abm_graph_launch(abm_graph *g, void *stream);
int abm_graph_destroy(abm_graph *g);
/* Stream-ordered device-to-device copy (capturable). */
int abm_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream);

/* ---- timing on the stream the kernels run on ---- */
int abm_event_timer(void *stream, int phase, float *ms_out);

#ifdef __cplusplus
}
#endif
#endif
