/* neuralce_hip.h — C ABI of libneuralce_hip.so, the gfx950 implementation of neuralce's parallel-tempering
 * canonical-swap Monte Carlo.
 *
 * The reference's boundary is a Python/JAX API, not an FFI; each entry point below replaces the JAX function named
 * in its comment (paths relative to the reference repository).  Conventions:
 *   - every call returns 0 on success or a negative error code; nce_last_error() describes the last failure of the
 *     calling thread;
 *   - buffers passed to compute calls are DEVICE pointers owned by the caller (torch tensors' data_ptr());
 *     the library owns only the context and the uploaded model tables;
 *   - compute calls are stream-ordered on `stream` (a hipStream_t, NULL = default stream) and do not synchronise;
 *   - one context per host thread; a context has no internal locking.
 * Sites are in canonical order i = cell * n_basis + basis, cell = (k0 d1 + k1) d2 + k2 (k = Smith-normal-form cell
 * coordinate), exactly as neuralce/ce/ce_gpu_implicit.py:1-21.
 */
#ifndef NEURALCE_HIP_H
#define NEURALCE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NCE_OK 0
#define NCE_ERR_ARG (-1)
#define NCE_ERR_HIP (-2)
#define NCE_ERR_UNSUPPORTED (-3)
#define NCE_ERR_NOMEM (-4)

#define NCE_PREC_FP64 0      /* plain form: FP64 swap energy change */
#define NCE_PREC_MIXED 1     /* FP32 swap energy change, FP64 running totals / exchange / accumulators */
#define NCE_PREC_BF16 2      /* NeuralCE only: as MIXED, with the graph-convolution projections on BF16 MFMA
                                (FP32 accumulation) */

typedef struct nce_ctx nce_ctx;

/* Cluster-expansion model on one cell shape: cell-independent template (reverse entries grouped into per-basis
 * windows, forward rows) plus the cell's reduced offsets.  Host arrays; copied to the device by nce_ce_model_create.
 * Replaces the `tp` pytree built by build_ce_gpu_implicit (neuralce/ce/ce_gpu_implicit.py:50-126). */
typedef struct {
    int32_t n_basis;          /* nB */
    int32_t max_order;        /* mo <= 4 */
    int32_t n_entries;        /* J: reverse entries per basis (padded) */
    int32_t max_window;       /* Wmax: window slots per basis (padded) */
    int32_t n_rows;           /* Rt: forward rows */
    int32_t n_sublattices;
    int32_t n_active;         /* active sublattices (proposal groups) */
    int32_t max_active_basis;
    int32_t d[3];             /* Smith normal form diagonal of the cell */
    int64_t n_cells;
    double factor;            /* energy_scale / n_cells */
    double zerolet;           /* zerolet * energy_scale */
    const int32_t *entries_per_basis;  /* [nB] */
    const int32_t *window_per_basis;   /* [nB] */
    const uint16_t *entry_slot;        /* [nB][J][4] window slot of each member; 0xFFFF = padding (value 1) */
    const double *entry_coef;          /* [nB][J] eci_orbit / n_instances_orbit; 0 for padding entries */
    const int32_t *window;             /* [nB][Wmax][4] (basis, k0, k1, k2): U*offset mod d */
    const int32_t *rows;               /* [Rt][mo][4] forward rows (basis, U*offset mod d), padding = anchor */
    const int32_t *row_bf;             /* [Rt][mo] point-function index (0 = constant 1) */
    const double *row_coef;            /* [Rt] */
    const double *pf;                  /* [n_sublattices][2][2] point functions pf[sl][bf][species] */
    const int32_t *basis_to_sl;        /* [nB] */
    const int32_t *active_basis;       /* [n_active][max_active_basis] */
    const int32_t *active_nbasis;      /* [n_active] */
    const double *active_cum;          /* [n_active] cumulative selection probability */
} nce_ce_model_desc;

/* One parallel-tempering job for nce_pt_run_chunk.  All pointers are DEVICE pointers. */
typedef struct {
    int32_t model;              /* id returned by nce_ce_model_create */
    int32_t n_replicas;         /* R */
    int64_t n_sites;            /* N (= n_cells * n_basis of the model) */
    int8_t *occ;                /* [R][N] species index per site, canonical order */
    double *energy;             /* [R] running total energies */
    const double *beta;         /* [R] 1/(kB T) of each replica's current temperature slot */
    const int32_t *slot_of_rep; /* [R] */
    int64_t *step;              /* [R] global step counter (Philox counter), advanced by n_steps */
    uint64_t seed;              /* Philox key of the job */
    int64_t *n_accept;          /* [R] accepted swaps, accumulated */
    double *e_min;              /* [R] running minimum energy */
    double *traj;               /* [n_steps / sample_interval][R] energy by slot after every interval, or NULL */
    int32_t sample_interval;
    int32_t accumulate;         /* 1: add (E - e_ref) and (E - e_ref)^2 of every step to s1, s2 */
    const double *e_ref;        /* [R] */
    double *s1, *s2;            /* [R] */
    const int64_t *inj_ia;      /* [R][n_steps] injected proposals (site a), or NULL for the Philox stream */
    const int64_t *inj_ib;      /* [R][n_steps] */
    const double *inj_u;        /* [R][n_steps] Metropolis uniforms */
    double *acc_n;              /* [R] by temperature slot, or NULL.  With accumulate: after the chunk, each replica's
                                   (s1, s2) around e_ref are folded into the FP64 moments of its slot (count, mean,
                                   sum of squared deviations; Chan's parallel merge), then s1 = s2 = 0 and
                                   e_ref = energy, on the device */
    double *acc_mean;           /* [R] */
    double *acc_m2;             /* [R] */
} nce_pt_job;

/* Temperature-swap exchange of one job (device pointers). */
typedef struct {
    int32_t n_replicas;
    int32_t inj_parity;         /* >= 0: injected parity, < 0: Philox */
    const double *energy;       /* [R] per replica */
    int32_t *rep_of_slot;       /* [R] updated */
    int32_t *slot_of_rep;       /* [R] updated */
    const double *beta_slot;    /* [R] ladder */
    const int32_t *ladder_id;   /* [R] or NULL */
    double *beta_rep;           /* [R] updated: beta of each replica's new slot */
    uint64_t seed;
    uint64_t ex_index;          /* exchange counter (Philox counter) */
    int64_t *n_swaps;           /* [1] accumulated accepted exchanges */
    const double *inj_u;        /* [ceil((R-1)/2)] injected uniforms or NULL */
    double *traj_last;          /* [R] if not NULL: rewritten as energy[rep_of_slot[t]] after the exchange */
    int64_t *ex_count;          /* [1] or NULL.  If set, ex_index is ignored: the exchange uses *ex_count + 1 and
                                   stores it back, so a captured cycle draws a new parity and uniforms each replay */
} nce_pt_exchange_job;

typedef struct nce_pt_cycle nce_pt_cycle;

const char *nce_last_error(void);

/* Context on device `device_ordinal` (replaces JAX's device selection; neuralce/mc/run.py:242-245). */
int nce_ctx_create(int32_t device_ordinal, nce_ctx **out);
void nce_ctx_destroy(nce_ctx *ctx);

/* Upload a CE model (template + cell); replaces GPUClusterExpansion's implicit branch
 * (neuralce/ce/classical_ce_gpu.py:514-535) / build_ce_gpu_implicit (neuralce/ce/ce_gpu_implicit.py:50-126). */
int nce_ce_model_create(nce_ctx *ctx, const nce_ce_model_desc *desc, int32_t *model_id);

/* Full energies of R configurations: E[r] = zerolet + factor * sum_cells sum_rows coef * prod pf.
 * Replaces energy_a (neuralce/ce/ce_gpu_implicit.py:147-163) and its lax.map over replicas
 * (neuralce/ce/pt_mcmc_ce.py:845-846). */
int nce_ce_energy(nce_ctx *ctx, int32_t model, const int8_t *occ, int32_t n_replicas, int64_t n_sites,
                  double *energy, void *stream);

/* nce_ce_energy without any allocation: no reference counterpart; for callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).  The per-block partial sums live
 * in a fixed per-GPU array of NCE_NOALLOC_PARTIALS doubles (8 MiB, held until the process ends); a call needs
 * n_replicas * ceil(n_cells / 256) of them and fails with NCE_ERR_ARG beyond that.  Every call on a GPU uses that
 * GPU's one array, so calls that overlap on one GPU (other streams, other contexts, other threads) race; results
 * are those of nce_ce_energy. */
#define NCE_NOALLOC_PARTIALS (1 << 20)
int nce_ce_energy_noalloc(nce_ctx *ctx, int32_t model, const int8_t *occ, int32_t n_replicas, int64_t n_sites,
                          double *energy, void *stream);

/* Swap energy change of explicit proposals (one per replica): dE[r] for swapping ia[r] and ib[r] of occ[r].
 * Replaces delta_energy_a (neuralce/ce/ce_gpu_implicit.py:181-191). */
int nce_ce_delta(nce_ctx *ctx, int32_t model, const int8_t *occ, int32_t n_replicas, int64_t n_sites,
                 const int64_t *ia, const int64_t *ib, double *dE, int32_t precision, void *stream);

/* n_steps swap-Metropolis steps of every replica of every job, one launch (persistent within the chunk).
 * Replaces run_chunk = lax.scan(micro) of vmapped delta mc_step (neuralce/ce/pt_mcmc_ce.py:554-566,765-818) with
 * make_fast_swap_implicit proposals (neuralce/ce/ce_gpu_implicit.py:194-211).
 * jobs: HOST array of n_jobs job descriptors (copied to the device stream-ordered).
 * tile: 0 = default.  Binary one-basis models run the speculative kernel: tile 4, 8 or 16 = different-species
 * swaps of a replica evaluated in parallel (one wave each; 16 falls back to 8 where the tables do not fit in LDS).
 * The default takes 16, 8 or 4 by replicas per CU (< 2, < 4, more), then halves it while the expected number of
 * overlapping window pairs in a group, W(W-1)/2 * 4 * window / sites of the smallest cell, exceeds 2.  Other
 * models run the general kernel: tile 1, 2, 4 or 8 = waves per step. */
int nce_pt_run_chunk(nce_ctx *ctx, const nce_pt_job *jobs, int32_t n_jobs, int64_t n_steps, int32_t precision,
                     int32_t tile, void *stream);

/* Temperature-swap exchange of each job; replaces exchange_slots (neuralce/ce/pt_mcmc_ce.py:748-760) and the
 * reordering of the last sample after it (:808-810).  jobs: HOST array. */
int nce_pt_exchange(nce_ctx *ctx, const nce_pt_exchange_job *jobs, int32_t n_jobs, void *stream);

/* One exchange cycle of a fixed set of jobs captured as a HIP graph: chunks_per_cycle chunks of n_steps (each
 * followed by the moment merge of the jobs that give acc_n), then the exchange of every job in ex_jobs.  Replaces
 * the jitted run_chunk, whose scan runs chunk_size steps and fires exchange_slots at the exchange interval
 * (neuralce/ce/pt_mcmc_ce.py:798-818): each nce_pt_cycle_launch replays the whole cycle n_cycles times with one
 * graph launch each and no host work in between.  The job and exchange tables are copied into device memory owned
 * by the cycle at creation (the caller's buffers they point to must outlive it); the exchange jobs must give
 * ex_count; injected streams are refused.  Launch parameters (kernel, tile, LDS) are fixed at creation as
 * nce_pt_run_chunk would choose them.  A cycle belongs to its context's device; destroy it before the context. */
int nce_pt_cycle_create(nce_ctx *ctx, const nce_pt_job *jobs, int32_t n_jobs, int64_t n_steps,
                        int32_t chunks_per_cycle, int32_t precision, int32_t tile,
                        const nce_pt_exchange_job *ex_jobs, int32_t n_ex_jobs, nce_pt_cycle **out);
int nce_pt_cycle_launch(nce_pt_cycle *cycle, int32_t n_cycles, void *stream);
void nce_pt_cycle_destroy(nce_pt_cycle *cycle);

/* ── Neural cluster expansion (receptive-field local update) ─────────────────────────────────────────────────
 * Model on one cell shape: layer-0 template (window of the sites sharing a cluster with a site, reverse entries with
 * their channel and coefficient), shell partner offsets, and the network.  One basis site, binary species.
 * Replaces setup_implicit / build_geometry / make_local_fns_v3 (neuralce/nce/mc_model.py:175-228,
 * neuralce/nce/gpu_layer0_implicit.py:17-118, neuralce/nce/local_update_implicit.py:19-197). Host arrays. */
typedef struct {
    int32_t d[3];
    int64_t n_cells;
    int32_t n_window;                  /* W0: layer-0 window slots (slot 0 = the site) */
    const int32_t *window;             /* [W0][4] (basis, k0, k1, k2) reduced mod d */
    int32_t n_entries;                 /* J0 */
    const uint16_t *entry_slot;        /* [J0][4] window slots of the members, 0xFFFF = padding */
    const int32_t *entry_col;          /* [J0] layer-0 channel (1 + orbit) */
    const double *entry_coef;          /* [J0] nA / (n_instances * n_cells * order) */
    int32_t n_shells;                  /* P */
    const int32_t *shell_count;        /* [P] partners per shell */
    const int32_t *partners;           /* [sum P][4] (basis, k0, k1, k2) reduced mod d, grouped by shell */
    int32_t c0;                        /* layer-0 channels (zerolet included) */
    const double *mu, *sd, *lin;       /* [c0] standardisation and layer-0 readout */
    int32_t n_layers;                  /* L */
    const int32_t *layer_cin, *layer_d;/* [L] */
    const double *w_self;              /* per layer [cin][d], concatenated */
    const double *w_shell;             /* per layer [P][cin][d] */
    const double *w_gate;              /* per layer [P][cin][d] or NULL */
    const double *b, *gb;              /* per layer [d] (gb NULL without gate) */
    int32_t activation;                /* 0: tanh (times the sigmoid gate if w_gate), 1: SiLU */
    const double *head_w;              /* [d_L] */
    double head_b;
    int32_t readout;                   /* 0: mean over sites (+ head_b), 1: sum over sites (+ head_b) */
    double de_scale;                   /* model energy change -> total eV (ys * N for the reference models) */
    int32_t ball_cap[8];               /* advisory; the library derives the receptive-field sizes from the window,
                                          the partners and the cell itself */
    int32_t cache_mode;                /* 0: each replica's cache keeps the standardised layer-0 rows h0s [N][c0]
                                          (the reference's 'h0' cache mode); 1: it keeps only the hidden layers'
                                          pre-activations (the reference's 'proj' mode, chosen there for wide layer-0
                                          models: neuralce/nce/mc_model.py:202-206, local_update_implicit.py:161-183).
                                          The step never reads h0s rows in either mode; a full recompute in mode 1
                                          builds them in a temporary buffer, a few replicas at a time */
} nce_nn_model_desc;

/* One NeuralCE job for nce_nn_run_chunk (device pointers). */
typedef struct {
    int32_t model;
    int32_t n_replicas;
    int64_t n_sites;
    int8_t *occ;                /* [R][N] */
    void *cache;                /* [R][nce_nn_cache_floats] FP32 (mixed) or FP64 caches: h0s, hidden layers, e */
    void *scratch;              /* [R][nce_nn_scratch_floats] same type */
    double *energy;             /* [R] running totals (model units, standardised) */
    const double *beta;         /* [R] 1/(kB T) */
    int64_t *step;              /* [R] Philox counter */
    uint64_t seed;
    int64_t *n_accept;          /* [R] */
    double *traj;               /* [n_steps/sample_interval][R] or NULL */
    int32_t sample_interval;
    int32_t pad;
    const int64_t *inj_ia, *inj_ib;   /* [R][n_steps] injected proposals or NULL */
    const double *inj_u;
} nce_nn_job;

/* Upload a model.  Besides the receptive-field sizes the library builds, for this cell, the slots and sorted
 * partner hits of two flips whose balls share no site (the same for every such pair); NCE_NN_FIXED_BALLS=0 in the
 * environment at this call leaves them out, and every step then builds its balls in the LDS hash table.
 * NCE_NN_LDS_MODE_MAX=k (0..3) keeps at most the first k per-step arrays (layer-0 change, sorted hits, projections)
 * in LDS and the rest in the scratch; every placement computes the same step (a test and diagnosis switch).
 * nce_nn_sizes: per-replica sizes (elements of the precision's type) of the cache and the scratch. */
int nce_nn_model_create(nce_ctx *ctx, const nce_nn_model_desc *desc, int32_t *model_id);
int nce_nn_sizes(nce_ctx *ctx, int32_t model, int64_t *cache_elems, int64_t *scratch_elems);

/* Full caches and energies of R configurations (h0_full + every layer + energy_of; neuralce/nce/
 * local_update_implicit.py:89-112); energy[r] = sum or mean of the site energies + head_b (model units). */
int nce_nn_full_cache(nce_ctx *ctx, int32_t model, const int8_t *occ, int32_t n_replicas, int64_t n_sites,
                      void *cache, double *energy, int32_t precision, void *stream);

/* n_steps swap steps of every replica (local_delta + Metropolis + commit: neuralce/nce/mcmc.py:130-199,
 * local_update_implicit.py:117-194); a same-species or same-site proposal is an accepted step with dE = 0 as in
 * the reference (mcmc.py:147,179).  jobs: HOST array. */
int nce_nn_run_chunk(nce_ctx *ctx, const nce_nn_job *jobs, int32_t n_jobs, int64_t n_steps, int32_t precision,
                     void *stream);

/* Energy change (model units, as the running totals) of one swap per replica, the proposals given in
 * inj_ia / inj_ib (inj_u is not used), evaluated by the local update WITHOUT committing: occupations, caches,
 * energies and counters are left as they are.  dE: DEVICE array [total replicas of all jobs, in job order]; a
 * same-site or same-species proposal gives 0.  Replaces local_delta (neuralce/nce/local_update_implicit.py:117-183,
 * setup()['local_delta'] of neuralce/nce/mc_model.py:221-228); commit is a forced-accept step (u = 0). */
int nce_nn_delta(nce_ctx *ctx, const nce_nn_job *jobs, int32_t n_jobs, double *dE, int32_t precision, void *stream);

/* Philox4x32-10 of the library (known-answer checks). */
void nce_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

#ifdef __cplusplus
}
#endif
#endif
