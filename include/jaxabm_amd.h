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
/* vmap over keys[n][2] of jax.random.gamma(key, a, (per,)) (kind ABM_GAMMA),
 * loggamma (ABM_LOGGAMMA) or beta(key, a, b, (per,)) (ABM_BETA), float32,
 * with the element parameters a[n][per] (and b[n][per]) -> out[n][per].
 * Replaces the reference examples' random.beta calls
 * (examples/models/advanced_economic_model.py:106,351,635,958,962). */
#define ABM_GAMMA 0
#define ABM_LOGGAMMA 1
#define ABM_BETA 2
int abm_gamma_each(const uint32_t *keys, int64_t n, int64_t per, const float *a,
                   const float *b, int kind, float *out, int mode, void *stream);
/* The float32 transcendental functions the draws and model kernels call
 * (log1p inside erf_inv of jax.random.normal; log, log1p, pow and exp of
 * jax.random.gamma / beta, examples/models/advanced_economic_model.py:106;
 * exp of the Deffuant sigmoid gate, jaxabm/analysis.py:489-531), written with
 * IEEE basic operations only (csrc/abm_math.h), elementwise:
 * out[i] = f(x[i]) or pow(x[i], y[i]) (y NULL unless ABM_MATH_POW).
 * abm_math_f64: the float64 exp and log they are built on. */
#define ABM_MATH_EXP 0
#define ABM_MATH_LOG 1
#define ABM_MATH_LOG1P 2
#define ABM_MATH_POW 3
int abm_math_f32(int fn, const float *x, const float *y, float *out, int64_t n, void *stream);
int abm_math_f64(int fn, const double *x, double *out, int64_t n, void *stream);
/* jax.random.permutation(key, n) (jaxabm/analysis.py:87); workspace from
 * abm_permutation_workspace(n). */
size_t abm_permutation_workspace(int64_t n);
int abm_permutation_rounds(int64_t n);
int abm_permutation(const uint32_t *key, int64_t n, int32_t *out, void *ws, size_t ws_bytes,
                    int mode, void *stream);
/* permutation(keys[s * key_stride .. +2], n) for s < count into out[count][n]
 * (all steps' pairings of a run in one batched sort). */
size_t abm_permutation_batch_workspace(int64_t n, int32_t count);
int abm_permutation_batch(const uint32_t *keys, int64_t key_stride, int32_t count, int64_t n,
                          int32_t *out, void *ws, size_t ws_bytes, int mode, void *stream);
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
/* The same Barabasi-Albert edge list built on the device (replaces the
 * add_edge loop of jaxabm/agentpy.py:559-582 that builds a Network):
 * src/dst device int32[abm_ba_num_edges(n, m)], key host uint32[2],
 * unresolved device uint32[1] scratch.  Rounds of one thread per node resolve
 * the preferential-attachment draws against the closed-form repeated-node
 * list; the host checks for completion every 8 rounds (synchronises the
 * stream). */
int abm_ba_edges_device(int64_t n, int32_t m, const uint32_t key[2], int32_t *src,
                        int32_t *dst, uint32_t *unresolved, void *stream);
/* Undirected CSR (both directions, rows ascending). */
int abm_csr_undirected(int64_t n, int64_t n_edges, const int32_t *src, const int32_t *dst,
                       int32_t *row_ptr, int32_t *col);
/* CSR of a Network built on the device from m edge-list entries (replaces
 * the edge list grown by Network.add_edge and the masked get_neighbors
 * queries, jaxabm/agentpy.py:559-615).  src, dst, row_ptr[n+1] and col are
 * device arrays; col holds up to m entries, 2m with ABM_CSR_SYMMETRIZE.
 * Flags: ABM_CSR_SYMMETRIZE adds (dst, src) after every entry (both
 * directions, as abm_csr_undirected); ABM_CSR_DEDUP drops repeated (row, col)
 * entries; ABM_CSR_INSERTION_ORDER keeps each row in entry order (otherwise
 * neighbours ascend; not with ABM_CSR_DEDUP).  *nnz (host) receives the entry
 * count, also in row_ptr[n].  Synchronises the stream (set-up, not for graph
 * capture); an endpoint outside [0, n) is an error. */
#define ABM_CSR_SYMMETRIZE 1
#define ABM_CSR_DEDUP 2
#define ABM_CSR_INSERTION_ORDER 4
size_t abm_csr_from_edges_workspace(int64_t m, int32_t flags);
int abm_csr_from_edges(int64_t n, int64_t m, const int32_t *src, const int32_t *dst,
                       int32_t flags, int32_t *row_ptr, int32_t *col, int64_t *nnz, void *ws,
                       size_t ws_bytes, void *stream);
/* Contiguous agent chunks of about `target` cost (cost(v) = node_cost +
 * degree(v), plus heavy_cost when degree(v) > heavy_deg) and at most
 * max_nodes agents, with boundaries that are multiples of `align` (except the
 * last); writes chunk_ptr[n_chunks+1] (may be NULL to count); returns
 * n_chunks or -1. */
int64_t abm_node_chunks(int64_t n, const int32_t *row_ptr, int32_t align, int64_t node_cost,
                        int32_t heavy_deg, int64_t heavy_cost, int64_t target,
                        int32_t max_nodes, int32_t *chunk_ptr);

/* ---- generic CSR gather / scatter (neighbour aggregation for Network
 *      models; replaces the masked queries of jaxabm/agentpy.py:584-615) ---- */
/* out[v] = sum over the CSR row of v of x[col[e]] (float64 accumulation). */
int abm_csr_gather_sum_f32(int64_t n, const int32_t *row_ptr, const int32_t *col, const float *x,
                           float *out, void *stream);
/* out[v] = number of entries e in row v with mask[col[e]] != 0. */
int abm_csr_count_u8(int64_t n, const int32_t *row_ptr, const int32_t *col, const uint8_t *mask,
                     int32_t *out, void *stream);
/* out[idx[i]] += val[i] for i < m (atomic; idx within out). */
int abm_scatter_add_f32(int64_t m, const int32_t *idx, const float *val, float *out,
                        void *stream);
int abm_scatter_add_i32(int64_t m, const int32_t *idx, const int32_t *val, int32_t *out,
                        void *stream);

/* ---- network SIR program: the env-phase callback of JaxModel.step
 *      (jaxabm/model.py:182-190) + metrics_fn counts (model.py:194-200) ----
 * Device state per instance: state uint8[n] (low 2 bits S=0/I=1/R=2, bit 2 =
 * row degree > 32; read public codes with abm_sir_states), updated in place;
 * mark uint16[n], neighbour-infection step marks that let a step skip S
 * agents with no infected neighbour; and two infected-set bitmaps
 * uint32[2][abm_sir_bitmap_words(n)] (buffer 0 holds the current one between
 * runs).  Alignment: state 8 bytes, mark and bitmaps 16 bytes.  `t`
 * arguments are absolute step indices. */
typedef struct abm_sir_graph {
  int64_t n;                /* agents, below 2^31 - 64 */
  int64_t nnz;              /* CSR entries (2 x undirected edges), below 2^31 - 64 */
  const int32_t *row_ptr;   /* device [n+1] */
  const int32_t *col;       /* device [nnz] */
  const int32_t *chunk_ptr; /* device [n_chunks+1]: abm_node_chunks with the
                               parameters of abm_sir_chunk_params */
  int64_t n_chunks;
} abm_sir_graph;
/* Chunking the step kernel expects: alignment, per-agent cost, heavy-row
 * degree and extra cost, chunks per block, the persistent block count
 * (target cost = total / (blocks * chunks_per_block)) and the largest chunk
 * in agents. */
int abm_sir_chunk_params(int32_t *align, int64_t *node_cost, int32_t *heavy_deg,
                         int64_t *heavy_cost, int32_t *chunks_per_block, int32_t *blocks,
                         int32_t *max_nodes);
/* Words of one infected-set bitmap (a multiple of 4). */
int64_t abm_sir_bitmap_words(int64_t n);
/* Slots per step: one step's `slots` block holds slots*3 ints (count slots
 * plus one flag slot; callers zero it and never read it directly). */
int abm_sir_slots(void);
/* Initial state at step t0: all S with heavy-row flags, agents perm[0..n0)
 * infected, marks and bitmap (buffer 0) built from it.  Also validates the
 * chunk table (read back once, synchronising the stream): 0 .. n ascending,
 * inner boundaries multiples of 64, chunks of at most 8192 agents; a graph
 * whose table fails is refused, and the step entry points below assume a
 * graph that abm_sir_seed accepted. */
int abm_sir_seed(const abm_sir_graph *g, const int32_t *perm, int64_t n0, uint8_t *state,
                 uint16_t *mark, uint32_t *bitmap, uint32_t t0, void *stream);
/* Runs `steps` synchronous steps.  keys: device [steps][4] = (k_inf, k_rec)
 * per step.  slots: device scratch [steps][abm_sir_slots()][3].  counts:
 * device [steps][3] S/I/R after each step. */
int abm_sir_run(const abm_sir_graph *g, uint8_t *state, uint16_t *mark, uint32_t *bitmaps,
                uint32_t t0, const uint32_t *keys, int32_t steps, float beta, float gamma,
                int mode, int32_t *slots, int32_t *counts, void *stream);
/* One step: reads bitmap bm_in, writes bm_out, updates state in place;
 * accumulates into a zeroed slots[abm_sir_slots()][3].  prev_slots: the
 * previous step's slots (lets a step with no infected agent skip its work),
 * or NULL. */
int abm_sir_step(const abm_sir_graph *g, uint8_t *state, uint16_t *mark, const uint32_t *bm_in,
                 uint32_t *bm_out, uint32_t t, const uint32_t *keys4, float beta, float gamma,
                 int mode, const int32_t *prev_slots, int32_t *slots, void *stream);
/* counts[t] = sum of the count slots of step t, for the slots of consecutive
 * steps of one run whose first step was computed (prev_slots NULL or live).
 * A step that found no infected agent leaves its slots zero and gets the
 * previous step's counts.  Writes a mark word into each step's flag slot. */
int abm_sir_reduce_counts(int32_t *slots, int32_t steps, int32_t *counts, void *stream);
/* out[i] = public S/I/R code of state[i]. */
int abm_sir_states(const uint8_t *state, int64_t n, uint8_t *out, void *stream);
/* ceil(p * 2^23): integer form of `uniform < p`. */
uint32_t abm_unit_threshold(float p);
int abm_fill_i32(int32_t *p, int64_t n, int32_t v, void *stream);

/* ---- wealth-exchange program: per-agent update of AgentCollection.update
 *      (jaxabm/agent.py:132-177) + env-phase exchange + record-time mean and
 *      Gini (examples/models/advanced_economic_model.py:1798-1811) ---- */
typedef struct abm_wealth_inst {
  int64_t n;                  /* agents, >= 2 */
  float *x, *y;               /* device [n] */
  int32_t *w;                 /* device [n] wealth */
  int32_t *inc;               /* device [n] scratch, zero on entry, left zero */
  int32_t *hist;              /* device [n+1] scratch, zero on entry, left zero */
  const uint32_t *keys;       /* device [steps][4] = (ck, uk) per step */
  int32_t steps;
  float *mean_out, *gini_out; /* device [steps] */
  uint32_t *draws;            /* device [steps][3][n] scratch for the draws (computed
                               * grid-wide before the run), or NULL: drawn in the run */
} abm_wealth_inst;
/* x, y, wealth from the collection's init key (device uint32[2]). */
int abm_wealth_init(const uint32_t *init_key, int64_t n, float *x, float *y, int32_t *w,
                    int mode, void *stream);
/* Runs every instance of the device array insts[count] for its steps, one
 * workgroup per instance.  max_n: the largest n in the table (0 if unknown).
 * With 0 < max_n <= 8000 every instance runs out of LDS (20 bytes per agent
 * of dynamic LDS per workgroup); otherwise, and for an instance above max_n,
 * on its global buffers.  With draws, every instance's draws are computed
 * first by a grid-wide kernel, and an instance of n < 1024 agents keeps its
 * state in registers.  Same results either way. */
int abm_wealth_run_batch(const abm_wealth_inst *insts, int32_t count, int64_t max_n, int mode,
                         void *stream);
/* One instance across the whole GPU, for large n: two launches per step over
 * all workgroups (each agent's draws computed and applied, exchange atomics,
 * with the previous step's mean and Gini recorded by workgroup 0; then the
 * exchange applied, the wealth histogram, total and maximum).  inst: the
 * descriptor in host memory (passed to the kernels by value); its draws
 * field is not used.  acc: device scratch uint32[ABM_WEALTH_GRID_ACC_WORDS]
 * (the total and 16 copies of the 64 smallest histogram bins, so that no
 * word takes an atomic from every workgroup).  Same results as
 * abm_wealth_run_batch. */
#define ABM_WEALTH_GRID_ACC_WORDS 1026
int abm_wealth_run_grid(const abm_wealth_inst *inst, uint32_t *acc, int mode, void *stream);

/* ---- Moore-8 neighbourhood reductions over a (rows, cols) row-major Grid
 *      array (the reference Grid stores only shape and periodic flag,
 *      jaxabm/agentpy.py:465-527; its models reduce neighbourhoods with jnp
 *      code).  periodic != 0 wraps both axes; otherwise offsets outside the
 *      grid are left out. ---- */
/* out[r][c] = sum of x over the 8 neighbours, float64 accumulation in the
 * fixed order (-1,-1) (-1,0) (-1,1) (0,-1) (0,1) (1,-1) (1,0) (1,1), rounded
 * once to float32. */
int abm_grid_moore_sum_f32(int64_t rows, int64_t cols, int32_t periodic, const float *x,
                           float *out, void *stream);
/* same[r][c] = neighbours with cells != 0 and == cells[r][c]; occupied[r][c]
 * = neighbours with cells != 0 (either output may be NULL). */
int abm_grid_moore_counts_i8(int64_t rows, int64_t cols, int32_t periodic, const int8_t *cells,
                             int32_t *same, int32_t *occupied, void *stream);

/* ---- Schelling program on a periodic Grid (jaxabm/agentpy.py:465-527
 *      rebuilt with a Moore-8 step) ---- */
typedef struct abm_schelling_grid {
  int32_t height, width;
  int64_t n_agents;      /* set by abm_schelling_init */
  int8_t *cells;         /* device [H*W]: 0 empty, 1 / 2 agent type */
  int32_t *id;           /* device [H*W]: agent on the cell or -1 */
  int32_t *claim;        /* device [H*W]: INT32_MAX when unclaimed */
  int32_t *pos;          /* device [H*W]: cell of agent k (first n_agents) */
  int8_t *type;          /* device [H*W]: type of agent k */
  int32_t *target;       /* device [H*W]: scratch */
  int32_t *empty;        /* device [H*W], 16-byte aligned: scratch (empty-cell bitmap) */
  int32_t *block_counts; /* device [abm_schelling_block_count(H*W)]: scratch (cells
                            per 1024-cell block: agents at init; empty cells
                            during a run, in 8 copies summed by the readers) */
} abm_schelling_grid;
/* Length in int32 of the block_counts scratch for a grid of n_cells cells:
 * 16 x ceil(n_cells / 1024). */
int64_t abm_schelling_block_count(int64_t n_cells);
/* Cells from uniform(key, (H*W,)) and the row-major agent numbering; sets
 * g->n_agents (synchronises the stream once). */
int abm_schelling_init(abm_schelling_grid *g, const uint32_t *key, int mode, int32_t *n_agents_out,
                       void *stream);
/* Runs `steps` steps; keys: device [steps][2] collection keys ck.  counters:
 * device scratch [steps][2]; happy_frac, moves: device [steps].  Grids of up
 * to 15,360,000 cells. */
int abm_schelling_run(const abm_schelling_grid *g, const uint32_t *keys, int32_t steps, int mode,
                      int32_t *counters, float *happy_frac, int32_t *moves, void *stream);
/* Continues a run: the steps of abm_schelling_run, for a grid whose cells
 * have not changed since the last abm_schelling_run / abm_schelling_run_from
 * call on it ended (that call left the grid's bitmap and block counts current,
 * so they are not rebuilt).  prev_counters: device [2], the counters of the
 * step before the first one here, which the first step's identity test reads
 * (a step after one that found every agent happy is the identity).  The step
 * loop is jaxabm/model.py:146-216 split over several calls. */
int abm_schelling_run_from(const abm_schelling_grid *g, const uint32_t *keys, int32_t steps,
                           int mode, const int32_t *prev_counters, int32_t *counters,
                           float *happy_frac, int32_t *moves, void *stream);
/* Records `steps` settled steps without launching them.  A step whose
 * previous step found every agent happy is the identity (nothing moves, so
 * every agent is happy again), and so is every step after it.  prev_counters:
 * device [2], the counters of the step before the first one recorded; it must
 * hold a happy count of g->n_agents.  Writes counters (n_agents, 0), happy
 * fraction 1 and 0 moves for each step, exactly what abm_schelling_run gives
 * for those steps; when prev_counters holds another count, counters get -1,
 * the happy fraction NaN and moves -1 instead.  Lets a caller that reads the
 * count back stop launching a run once its grid has settled (the step loop is
 * jaxabm/model.py:146-216, which has no such exit). */
int abm_schelling_settled(const abm_schelling_grid *g, const int32_t *prev_counters, int32_t steps,
                          int32_t *counters, float *happy_frac, int32_t *moves, void *stream);

/* ---- differentiable Deffuant model, batch of (eps, mu) instances
 *      (loss and gradient set-up jaxabm/analysis.py:489-531; step schedule
 *      jaxabm/model.py:146-216) ---- */
typedef struct abm_deffuant_problem {
  int64_t n;               /* agents */
  int32_t batch;           /* instances B */
  int32_t steps;
  int32_t ckpt_every;      /* checkpoint interval k */
  int32_t pair_blocks;     /* workgroup rows of the step grids (<= 65535) */
  int32_t stat_chunks;     /* row chunks of the statistics pass (<= 65535) */
  int32_t flags;           /* ABM_DEFFUANT_* bits below */
  const int32_t *perms;    /* device [steps][n]: pairing of each step */
  const float *eps, *mu;   /* device [B] */
  float *work;             /* device [n][B]: final state after the forward */
  float *ckpt;             /* device [ceil(steps/k)][n][B]; ckpt[0] = initial state */
  float *traj;             /* device [k-1][n][B] (backward) */
  float *lam;              /* device [n][B] (backward) */
  double *acc;             /* device [pair_blocks][2][B] (backward) */
  double *partial;         /* device [stat_chunks][34][B] */
  double *sums;            /* device [34][B] */
  double *coef;            /* device [34][B] */
  const double *target;    /* device [34]: target [mean, var, hist_0..31] */
  double *stats;           /* device [B][34] */
  double *loss;            /* device [B] */
  double *grads;           /* device [2][B]: dL/deps, dL/dmu (backward) */
  const double *weights;   /* device [34] per-statistic loss weights, or NULL (all 1) */
} abm_deffuant_problem;
/* flags bit: in abm_deffuant_forward, also keep the last checkpoint segment's
 * intermediate states in traj (same bytes: those steps write out of place);
 * in abm_deffuant_backward, traj holds them from the forward just before on
 * this descriptor, so that segment is not recomputed.  A second backward of
 * the same forward must clear it (the first one overwrites traj). */
#define ABM_DEFFUANT_TAPE_LAST 1
/* x[i][b] = uniform(split(init_key, n)[i]) for every instance b. */
int abm_deffuant_init(const uint32_t *init_key, int64_t n, int32_t batch, float *x, int mode,
                      void *stream);
/* One step of the whole batch with pairing perm[n]; src == dst in place. */
int abm_deffuant_step(const float *src, float *dst, int64_t n, int32_t batch,
                      const int32_t *perm, const float *eps, const float *mu,
                      int32_t pair_blocks, void *stream);
/* Steps from ckpt[0], keeping the checkpoints (and, with
 * ABM_DEFFUANT_TAPE_LAST and traj, the last segment's states); then stats,
 * loss and the dL/dx coefficients of the final state. */
int abm_deffuant_forward(const abm_deffuant_problem *p, void *stream);
/* Reverse sweep after abm_deffuant_forward on the same descriptor:
 * recomputes each checkpoint segment, then grads = dL/d(eps, mu). */
int abm_deffuant_backward(const abm_deffuant_problem *p, void *stream);

/* ---- continuous-batching simulation server (replaces the serial runners
 *      jaxabm/analysis.py:97-165, jaxabm/utils.py:128-176,
 *      jaxabm/agentpy.py:1116-1152) ---- */
#define ABM_SERVE_WEALTH 0
#define ABM_SERVE_SIR 1
#define ABM_SERVE_QUEUED 0
#define ABM_SERVE_PREPARED 1
#define ABM_SERVE_UPLOADING 2
#define ABM_SERVE_ACTIVE 3
#define ABM_SERVE_DONE 4
#define ABM_SERVE_FAILED 5
typedef struct abm_serve_config {
  uint64_t hbm_budget;        /* device bytes for instance state (one arena) */
  int64_t max_agents;         /* largest request accepted */
  int32_t max_active;         /* instances advanced together (1: one at a time) */
  int32_t prefetch;           /* extra admitted instances uploading meanwhile */
  int32_t max_inflight;       /* ticks queued on the device (1..3) */
  int32_t steps_per_tick;     /* K */
  int32_t prep_threads;       /* host preparation workers */
  int32_t mode;               /* RNG stream mode */
  int64_t sir_chunk;          /* stepwise SIR: cost (4 + degree per agent) per range */
  int64_t wealth_chunk;       /* stepwise wealth: agents per range */
  int64_t sir_resident_max;   /* n <= this (and <= 65536): resident SIR instance */
  int64_t wealth_resident_max;/* n <= this (and <= 65536): resident wealth instance */
  int32_t strict_fifo;        /* 1: no admission past a request that does not fit */
  int32_t reserved;
} abm_serve_config;
typedef struct abm_request {
  int32_t kind;               /* ABM_SERVE_WEALTH or ABM_SERVE_SIR */
  int32_t steps;
  int64_t n;                  /* agents */
  int64_t seed;               /* model seed: PRNGKey(seed) */
  int64_t graph_seed;         /* SIR: Barabasi-Albert graph seed */
  int32_t m;                  /* SIR: edges per new node */
  int32_t n_infected;         /* SIR: initial infections */
  float beta, gamma;          /* SIR: infection / recovery probabilities */
} abm_request;
typedef struct abm_ticket_info {
  int32_t status, kind, steps, resident;
  int64_t n, nnz, device_bytes, result_bytes;
  double t_submit, t_ready, t_admit, t_start, t_done; /* abm_serve_clock() seconds */
} abm_ticket_info;
typedef struct abm_serve_stats {
  int64_t ticks, completed, agent_steps, peak_active, peak_device_bytes, outstanding;
} abm_serve_stats;
void abm_serve_default_config(abm_serve_config *cfg);
double abm_serve_clock(void);
int abm_serve_create(const abm_serve_config *cfg, void **server);
int abm_serve_submit(void *server, const abm_request *req, int64_t *ticket);
int abm_serve_poll(void *server, int64_t ticket, int32_t *status);
int abm_serve_wait(void *server, int64_t ticket, double timeout_s, int32_t *status);
/* Result records: SIR int32 counts[steps][3] (S, I, R); wealth float32
 * mean[steps] then gini[steps].  out may be NULL to read only the info. */
int abm_serve_collect(void *server, int64_t ticket, void *out, size_t out_bytes,
                      abm_ticket_info *info);
int abm_serve_drain(void *server, double timeout_s);
int abm_serve_get_stats(void *server, abm_serve_stats *st);
int abm_serve_destroy(void *server);

/* ---- step-loop capture (replaces the Python loop of JaxModel.run,
 *      jaxabm/model.py:240-241) ---- */
typedef struct abm_graph abm_graph;
int abm_graph_begin(void *stream);
int abm_graph_end(void *stream, abm_graph **out);
int abm_graph_launch(abm_graph *g, void *stream);
int abm_graph_destroy(abm_graph *g);
/* Prepares a captured, not yet instantiated hipGraph_t (a generic model step
 * of torch ops; replaces the per-step dispatch of jaxabm/model.py:146-216)
 * for replay: every memset node becomes a fill-kernel node with the same
 * edges; a graph holding host, event, allocation, symbol-copy or child-graph
 * nodes is refused (kErrInvalid) and left unchanged.  counts[4]: memsets
 * replaced, kernel nodes, memcpy nodes, empty nodes. */
int abm_graph_sanitize(void *graph, int32_t *counts);
/* One launch of up to ABM_MAX_COPY_ITEMS device copies: item i copies
 * `bytes` from src to dst + row[0] * row_stride (row_stride 0: dst as is;
 * rows at or past row_limit are skipped).  The items (a host array) travel
 * in the kernel arguments, so the launch can be captured into a graph.  The
 * per-step result copies and history writes of JaxModel.run(graph=True)
 * (jaxabm/model.py:146-216: the new states, env and metrics of a step). */
#define ABM_MAX_COPY_ITEMS 64
typedef struct {
  const void *src;
  void *dst;
  int64_t bytes;
  int64_t row_stride;
  int64_t row_limit;
} abm_copy_item;
int abm_copy_items(const abm_copy_item *items, int32_t n_items, const int64_t *row, void *stream);
/* row[0] += 1 and ts[0] += 1 (the replayed step's history row and time step). */
int abm_step_counters(int64_t *row, int32_t *ts, void *stream);
/* abm_copy_items plus the step's tail in the same launch: the last block to
 * finish advances row[0] and ts[0], returns *done (a device counter, 0
 * between launches) to 0, and copies key_words words of row row[0] of
 * key_table[key_rows][key_words] to key_dst (when the new row is below
 * key_rows; key_dst may be NULL): the next step's keys.  One node per
 * replayed step of JaxModel.run(graph=True) for all of its bookkeeping. */
int abm_commit_step(const abm_copy_item *items, int32_t n_items, int64_t *row, int32_t *ts,
                    uint32_t *done, const uint32_t *key_table, int64_t key_rows,
                    int32_t key_words, uint32_t *key_dst, void *stream);
/* Stream-ordered device-to-device copy (capturable). */
int abm_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream);

/* ---- record-time reductions over one agent variable (jaxabm/model.py:194-200,
 *      jaxabm/agentpy.py:926-942, jaxabm/agent.py:198-211; Gini as
 *      examples/models/advanced_economic_model.py:1798-1811).  Device inputs
 *      and outputs, no host synchronisation, no memset (graph-capturable). ---- */
/* counts[v] = number of states[i] == v for v < k (1 <= k <= 256), exact;
 * states may be NULL when n == 0. */
int abm_reduce_counts_u8(const uint8_t *states, int64_t n, int32_t k, int32_t *counts,
                         void *stream);
/* Scratch bytes for any reduction below over n values. */
size_t abm_reduce_workspace(int64_t n);
/* *out = mean: float32 input summed in float64 in a fixed order, rounded once;
 * int32 input summed exactly, then f32(sum) / f32(n) (jnp.mean of int32). */
int abm_mean_f32(const float *x, int64_t n, float *out, void *ws, size_t ws_bytes,
                 void *stream);
int abm_mean_i32(const int32_t *x, int64_t n, float *out, void *ws, size_t ws_bytes,
                 void *stream);
/* Whole-array sums (jnp.sum with x64 off inside metrics / env callbacks):
 * float32 summed in float64 in a fixed order and rounded once; int32 and
 * uint8 (booleans) summed exactly into int32 with wrap-around.  Up to 65,536
 * values one launch and no workspace (ws may be NULL); more need
 * abm_reduce_workspace(n) bytes.  The means above take the same one-launch
 * path up to 65,536 values. */
int abm_sum_f32(const float *x, int64_t n, float *out, void *ws, size_t ws_bytes, void *stream);
int abm_sum_i32(const int32_t *x, int64_t n, int32_t *out, void *ws, size_t ws_bytes,
                void *stream);
int abm_sum_u8(const uint8_t *x, int64_t n, int32_t *out, void *ws, size_t ws_bytes,
               void *stream);
/* Several of the sums and means above in one launch, each of up to 65,536
 * values and computed exactly as its own call computes it (same order, same
 * rounding): the reductions of one step's env and metrics callbacks
 * (jaxabm/model.py:182-200 update_state_fn / metrics_fn, jaxabm/agentpy.py:
 * 926-942 compute_metrics), batched.  op: ABM_REDUCE_SUM_F32 (float *out),
 * ABM_REDUCE_MEAN_F32, ABM_REDUCE_SUM_I32 (int32_t *out), ABM_REDUCE_MEAN_I32
 * (float), ABM_REDUCE_SUM_U8 (int32_t).  1 <= count <= ABM_REDUCE_BATCH_MAX. */
#define ABM_REDUCE_BATCH_MAX 16
#define ABM_REDUCE_SUM_F32 0
#define ABM_REDUCE_MEAN_F32 1
#define ABM_REDUCE_SUM_I32 2
#define ABM_REDUCE_MEAN_I32 3
#define ABM_REDUCE_SUM_U8 4
typedef struct abm_reduce_item {
  const void *x;
  int64_t n;
  void *out;
  int32_t op;
  int32_t reserved;
} abm_reduce_item;
int abm_reduce_batch(const abm_reduce_item *items, int32_t count, void *stream);
/* *out = Gini of int32 values: sorted x, S = sum (i+1) x_(i), T = sum x in
 * int32 (wrapping), G = f32(2S) / f32(nT) - f32((n+1)/n), 0 when T <= 1e-6. */
int abm_gini_i32(const int32_t *x, int64_t n, float *out, void *ws, size_t ws_bytes,
                 void *stream);
/* abm_gini_i32 without a workspace and without any allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Its scratch is a 5.25 MiB device array
 * of the library, one copy per GPU held until the process ends, so n is at
 * most ABM_GINI_NOALLOC_MAX_N (larger n is refused with an error).  Every
 * call on a GPU uses that GPU's one copy: calls that overlap on one GPU (two
 * streams, two threads) race and must be ordered by the caller. */
#define ABM_GINI_NOALLOC_MAX_N 262144
int abm_gini_i32_noalloc(const int32_t *x, int64_t n, float *out, void *stream);
/* out[k] (float64, k < bins <= 256) = mean over i of
 * sigmoid(slope (x_i - k/bins)) - sigmoid(slope (x_i - (k+1)/bins)). */
int abm_soft_hist_f32(const float *x, int64_t n, int32_t bins, double slope, double *out,
                      void *ws, size_t ws_bytes, void *stream);

/* ---- generated agent-update kernels (replace the jax.vmap of a per-agent
 *      function in AgentCollection.update, jaxabm/agent.py:156-177, which XLA
 *      fuses) ---- */
/* Compiles kernel source text for `arch` (e.g. "gfx950") with hiprtc; *code
 * (free with abm_rtc_code_free) receives the code object.  Needs no GPU. */
int abm_rtc_compile(const char *src, const char *arch, void **code, size_t *size);
int abm_rtc_code_free(void *code);
/* Loads a code object and finds `kernel` in it. */
int abm_rtc_load(const void *code, const char *kernel, void **handle);
int abm_rtc_unload(void *handle);
/* Launches a loaded kernel whose one parameter is a struct of args_bytes
 * bytes (64-bit words) on `stream`. */
int abm_rtc_launch(void *handle, uint32_t grid, uint32_t block, const void *args,
                   size_t args_bytes, void *stream);

/* ---- timing on the stream the kernels run on ---- */
int abm_event_timer(void *stream, int phase, float *ms_out);

#ifdef __cplusplus
}
#endif
#endif
