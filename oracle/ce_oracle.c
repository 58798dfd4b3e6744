/* FP64 CPU oracle for the cluster-expansion parallel-tempering swap Monte Carlo.
 *
 * TEST INFRASTRUCTURE ONLY: this file is the checker the HIP kernels are compared against, and the CPU baseline
 * timed beside them.  Nothing in the product links or calls it.
 *
 * It restates the reference's JAX functions in plain C, in the reference's own formulation (per-entry member site
 * lists with species overrides), deliberately NOT the window / value-table formulation of the GPU kernel:
 *   oracle_ce_energy     neuralce/ce/ce_gpu_implicit.py:147-163 (energy_a: forward rows over every cell)
 *   oracle_ce_delta      neuralce/ce/ce_gpu_implicit.py:166-191 (_entries, _species_at, delta_energy_a: entries of
 *                        i_a with both overrides, plus entries of i_b that do not contain i_a)
 *   oracle_pt_chunk      neuralce/ce/pt_mcmc_ce.py:554-566 (delta mc_step: valid = species differ, accept =
 *                        valid && log u < -beta dE, two-site write, E += dE) and :765-818 (chunk of steps, E_min,
 *                        subsampling at the (si-1)-th step of every interval)
 *   oracle_pt_exchange   neuralce/ce/pt_mcmc_ce.py:748-760 (exchange_slots: random parity, neighbour slot pairs of
 *                        one ladder, log u < (b_j - b_i)(E_j - E_i), permutation of rep_of_slot, beta per replica)
 * Proposals (ce_gpu_implicit.py:194-211) are drawn either from injected arrays (parity tests) or from the same
 * counter-based Philox4x32-10 stream definition the kernels use (Salmon et al., SC'11), so a Philox-driven GPU chunk can
 * also be replayed exactly.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    int32_t nB, mo, J, Rt, n_sl;
    const int32_t *T_b, *T_uo, *T_bf;      /* (Rt,mo), (Rt,mo,3) reduced mod d, (Rt,mo) */
    const double *T_coef;                  /* (Rt) eci / n_instances */
    const int32_t *Rv_b, *Rv_uo, *Rv_bf;   /* (nB,J,mo), (nB,J,mo,3), (nB,J,mo) */
    const double *Rv_coef;                 /* (nB,J); 0 for padded entries */
    const double *pf;                      /* (n_sl,2,2) */
    const int32_t *basis_to_sl;            /* (nB) */
    int32_t d[3];
    int64_t n_cells;
    double factor, zerolet;                /* energy_scale / n_cells, zerolet * energy_scale */
    int32_t n_act, maxb;
    const int32_t *act_b, *act_nb;         /* (n_act,maxb), (n_act) */
    const double *act_cum;                 /* (n_act) cumulative selection probability */
} oracle_ce;

/* ── Philox4x32-10 ─────────────────────────────────────────────────────────────────────────────────── */
static inline void mulhilo32(uint32_t a, uint32_t b, uint32_t *hi, uint32_t *lo) {
    uint64_t p = (uint64_t)a * (uint64_t)b; *hi = (uint32_t)(p >> 32); *lo = (uint32_t)p;
}

void oracle_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo32(0xD2511F53u, c0, &hi0, &lo0);
        mulhilo32(0xCD9E8D57u, c2, &hi1, &lo1);
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline uint64_t umulhi64(uint64_t a, uint64_t b) { return (uint64_t)(((unsigned __int128)a * b) >> 64); }

/* proposal + Metropolis uniform of one step (stream layout shared with the kernels):
 * x = philox({step_lo, step_hi, replica, 0}, seed), y = philox({..., 1}, seed)
 * cell_a = hi64(x1:x0 * n_cells), cell_b = hi64(x3:x2 * n_cells); sublattice by y2 * 2^-32 against the cumulative
 * probabilities; basis index (y3 & 0xffff) * nb >> 16 and (y3 >> 16) * nb >> 16; u = ((y1:y0 >> 11) + 0.5) 2^-53. */
void oracle_draw(const oracle_ce *m, uint64_t seed, uint64_t step, uint32_t replica, int64_t *ia, int64_t *ib, double *u) {
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t ctr[4] = {(uint32_t)step, (uint32_t)(step >> 32), replica, 0u};
    uint32_t x[4], y[4];
    oracle_philox4x32_10(ctr, key, x);
    ctr[3] = 1u;
    oracle_philox4x32_10(ctr, key, y);
    uint64_t ca = umulhi64(((uint64_t)x[1] << 32) | x[0], (uint64_t)m->n_cells);
    uint64_t cb = umulhi64(((uint64_t)x[3] << 32) | x[2], (uint64_t)m->n_cells);
    int sl = 0;
    if (m->n_act > 1) {
        double us = (double)y[2] * (1.0 / 4294967296.0);
        while (sl < m->n_act - 1 && !(us < m->act_cum[sl])) ++sl;
    }
    uint32_t nb = (uint32_t)m->act_nb[sl];
    int ba = m->act_b[sl * m->maxb + (int)(((y[3] & 0xffffu) * nb) >> 16)];
    int bb = m->act_b[sl * m->maxb + (int)(((y[3] >> 16) * nb) >> 16)];
    *ia = (int64_t)ca * m->nB + ba;
    *ib = (int64_t)cb * m->nB + bb;
    *u = ((double)((((uint64_t)y[1] << 32) | y[0]) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

/* ── geometry ──────────────────────────────────────────────────────────────────────────────────────── */
static inline int64_t site_of(const oracle_ce *m, int32_t b, int64_t k0, int64_t k1, int64_t k2) {
    return ((k0 * m->d[1] + k1) * m->d[2] + k2) * m->nB + b;
}

static inline double pfv(const oracle_ce *m, int32_t b, int32_t bf, int sp) {
    return m->pf[(m->basis_to_sl[b] * 2 + bf) * 2 + sp];
}

double oracle_ce_energy(const oracle_ce *m, const int8_t *occ) {
    const int64_t n_cells = m->n_cells;
    double total = 0.0;
    for (int64_t c = 0; c < n_cells; ++c) {
        int64_t k0 = c / ((int64_t)m->d[1] * m->d[2]), k1 = (c / m->d[2]) % m->d[1], k2 = c % m->d[2];
        double cs = 0.0;
        for (int r = 0; r < m->Rt; ++r) {
            double p = 1.0;
            for (int q = 0; q < m->mo; ++q) {
                const int32_t *uo = m->T_uo + ((int64_t)r * m->mo + q) * 3;
                int32_t b = m->T_b[r * m->mo + q];
                int64_t s = site_of(m, b, (k0 + uo[0]) % m->d[0], (k1 + uo[1]) % m->d[1], (k2 + uo[2]) % m->d[2]);
                p *= pfv(m, b, m->T_bf[r * m->mo + q], occ[s]);
            }
            cs += m->T_coef[r] * p;
        }
        total += cs;
    }
    return m->zerolet + m->factor * total;
}

/* entries of site i: member sites (J,mo) */
static void entries(const oracle_ce *m, int64_t i, int64_t *sites, int *bj) {
    int32_t b = (int32_t)(i % m->nB); int64_t c = i / m->nB;
    int64_t k0 = c / ((int64_t)m->d[1] * m->d[2]), k1 = (c / m->d[2]) % m->d[1], k2 = c % m->d[2];
    *bj = b;
    for (int j = 0; j < m->J; ++j)
        for (int q = 0; q < m->mo; ++q) {
            const int32_t *uo = m->Rv_uo + (((int64_t)b * m->J + j) * m->mo + q) * 3;
            int32_t bb = m->Rv_b[((int64_t)b * m->J + j) * m->mo + q];
            sites[j * m->mo + q] = site_of(m, bb, (k0 + uo[0]) % m->d[0], (k1 + uo[1]) % m->d[1], (k2 + uo[2]) % m->d[2]);
        }
}

double oracle_ce_delta(const oracle_ce *m, const int8_t *occ, int64_t ia, int64_t ib, int sa_new, int sb_new) {
    int64_t *sites = (int64_t *)malloc(sizeof(int64_t) * (size_t)m->J * m->mo);
    double dA = 0.0, dB = 0.0;
    for (int pass = 0; pass < 2; ++pass) {
        int b;
        entries(m, pass == 0 ? ia : ib, sites, &b);
        for (int j = 0; j < m->J; ++j) {
            const int64_t *sj = sites + (int64_t)j * m->mo;
            if (pass == 1) {                                    /* i_b entries that also contain i_a are skipped */
                int both = 0;
                for (int q = 0; q < m->mo; ++q) both |= (sj[q] == ia);
                if (both) continue;
            }
            double po = 1.0, pn = 1.0;
            for (int q = 0; q < m->mo; ++q) {
                int64_t s = sj[q];
                int32_t bb = m->Rv_b[((int64_t)b * m->J + j) * m->mo + q];
                int32_t bf = m->Rv_bf[((int64_t)b * m->J + j) * m->mo + q];
                int sp = occ[s], spn = sp;
                if (s == ia) spn = sa_new;                      /* overrides in order: (i_a, new), (i_b, new) */
                if (s == ib) spn = sb_new;
                po *= pfv(m, bb, bf, sp);
                pn *= pfv(m, bb, bf, spn);
            }
            double c = m->Rv_coef[(int64_t)b * m->J + j];
            if (pass == 0) dA += c * (pn - po); else dB += c * (pn - po);
        }
    }
    free(sites);
    return m->factor * (dA + dB);
}

/* ── chunk of PT steps ─────────────────────────────────────────────────────────────────────────────── */
typedef struct {
    int32_t R;                     /* replicas */
    int64_t N;                     /* sites per replica */
    int8_t *occ;                   /* [R][N] */
    double *E;                     /* [R] running totals */
    const double *beta;            /* [R] beta of the replica's current slot */
    const int32_t *slot_of_rep;    /* [R] */
    uint64_t seed, step0;          /* Philox key, global step of the first step of the chunk */
    int64_t *n_acc;                /* [R] accepted swaps (accumulated) */
    double *e_min;                 /* [R] running minimum (updated) */
    double *traj;                  /* [n_steps/si][R] by slot, may be NULL */
    int32_t sample_interval;
    int32_t accumulate;            /* add (E - e_ref) and (E - e_ref)^2 of every step to s1, s2 */
    const double *e_ref;           /* [R] */
    double *s1, *s2;               /* [R] */
    const int64_t *inj_ia, *inj_ib;  /* [R][n_steps] injected proposals (or NULL: Philox) */
    const double *inj_u;           /* [R][n_steps] */
} oracle_chunk;

int oracle_pt_chunk(const oracle_ce *m, const oracle_chunk *s, int64_t n_steps, int n_threads) {
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int r = 0; r < s->R; ++r) {
        int8_t *occ = s->occ + (int64_t)r * s->N;
        double E = s->E[r], beta = s->beta[r], emin = s->e_min[r];
        double s1 = 0.0, s2 = 0.0; int64_t acc = 0;
        double eref = s->accumulate ? s->e_ref[r] : 0.0;
        for (int64_t t = 0; t < n_steps; ++t) {
            int64_t ia, ib; double u;
            if (s->inj_ia) {
                ia = s->inj_ia[(int64_t)r * n_steps + t]; ib = s->inj_ib[(int64_t)r * n_steps + t];
                u = s->inj_u[(int64_t)r * n_steps + t];
            } else {
                oracle_draw(m, s->seed, s->step0 + (uint64_t)t, (uint32_t)r, &ia, &ib, &u);
            }
            int sa = occ[ia], sb = occ[ib];
            if (sa != sb) {
                double dE = oracle_ce_delta(m, occ, ia, ib, sb, sa);
                if (log(u) < -beta * dE) {
                    occ[ia] = (int8_t)sb; occ[ib] = (int8_t)sa; E += dE; ++acc;
                }
            }
            if (E < emin) emin = E;
            if (s->accumulate) { double x = E - eref; s1 += x; s2 += x * x; }
            if (s->traj && ((t + 1) % s->sample_interval) == 0)
                s->traj[((t + 1) / s->sample_interval - 1) * s->R + s->slot_of_rep[r]] = E;
        }
        s->E[r] = E; s->e_min[r] = emin; s->n_acc[r] += acc;
        if (s->accumulate) { s->s1[r] += s1; s->s2[r] += s2; }
    }
    return 0;
}

/* ── temperature-swap exchange ─────────────────────────────────────────────────────────────────────── */
/* parity = z0 & 1 of philox({ex_lo, ex_hi, 0xffffffff, 2}); pair p uses u = ((w1:w0 >> 11) + 0.5) 2^-53 of
 * philox({ex_lo, ex_hi, p, 3}).  inj_parity >= 0 and inj_u select injected values instead. */
int oracle_pt_exchange(int32_t R, const double *E, int32_t *rep_of_slot, const double *beta_slot,
                       const int32_t *ladder_id, double *beta_rep, uint64_t seed, uint64_t ex_index,
                       int32_t inj_parity, const double *inj_u, int64_t *n_swaps) {
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    int parity;
    if (inj_parity >= 0) parity = inj_parity & 1;
    else {
        uint32_t ctr[4] = {(uint32_t)ex_index, (uint32_t)(ex_index >> 32), 0xffffffffu, 2u}, z[4];
        oracle_philox4x32_10(ctr, key, z); parity = (int)(z[0] & 1u);
    }
    int64_t ns = 0;
    for (int32_t p = 0;; ++p) {
        int32_t i = 2 * p + parity, j = i + 1;
        if (2 * p > R - 2) break;                              /* i_idx = arange(0, R-1, 2) + parity */
        if (j >= R) continue;
        if (ladder_id && ladder_id[i] != ladder_id[j]) continue;
        double u;
        if (inj_u) u = inj_u[p];
        else {
            uint32_t ctr[4] = {(uint32_t)ex_index, (uint32_t)(ex_index >> 32), (uint32_t)p, 3u}, w[4];
            oracle_philox4x32_10(ctr, key, w);
            u = ((double)((((uint64_t)w[1] << 32) | w[0]) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        }
        double la = (beta_slot[j] - beta_slot[i]) * (E[rep_of_slot[j]] - E[rep_of_slot[i]]);
        if (log(u) < la) { int32_t t = rep_of_slot[i]; rep_of_slot[i] = rep_of_slot[j]; rep_of_slot[j] = t; ++ns; }
    }
    for (int32_t t = 0; t < R; ++t) beta_rep[rep_of_slot[t]] = beta_slot[t];
    *n_swaps += ns;
    return 0;
}

/* ── exact enumeration ────────────────────────────────────────────────────────────────────────────────────────
 * Energies of every configuration of n binary sites with k of species 1, in lexicographic order of the k-subsets:
 * E = e0 + sum_t coef[t] prod_{m < mo, site[t][m] >= 0} s[site[t][m]], s = +1 for species 1 and -1 for species 0.
 * The sequence is cut into ranges, one per thread; each range starts from its subset unranked from the
 * combinatorial number system and steps with the next-subset rule.  n <= 62. */
static double binom(int n, int k) {
    if (k < 0 || k > n) return 0.0;
    double r = 1.0;
    for (int i = 1; i <= k; ++i) r = r * (double)(n - k + i) / (double)i;
    return r;
}

static void unrank_subset(int n, int k, int64_t rank, int *c) {
    int x = 0;
    for (int i = 0; i < k; ++i) {
        for (;; ++x) {
            const int64_t cnt = (int64_t)llround(binom(n - x - 1, k - i - 1));    /* subsets starting with x */
            if (rank < cnt) break;
            rank -= cnt;
        }
        c[i] = x++;
    }
}

int oracle_enumerate_energies(int32_t n, int32_t k, int64_t n_terms, int32_t mo, const int32_t *site,
                              const double *coef, double e0, double *E_out) {
    if (n < 1 || n > 62 || k < 0 || k > n || mo < 1) return -1;
    const int64_t total = (int64_t)llround(binom(n, k));
    int n_threads = 1;
#ifdef _OPENMP
    n_threads = omp_get_max_threads();
#endif
#pragma omp parallel for schedule(static, 1)
    for (int t = 0; t < n_threads; ++t) {
        const int64_t lo = total * t / n_threads, hi = total * (t + 1) / n_threads;
        int c[64];
        double s[64];
        if (lo < hi) {
            unrank_subset(n, k, lo, c);
            for (int64_t r = lo; r < hi; ++r) {
                for (int i = 0; i < n; ++i) s[i] = -1.0;
                for (int i = 0; i < k; ++i) s[c[i]] = 1.0;
                double e = e0;
                for (int64_t q = 0; q < n_terms; ++q) {
                    double p = coef[q];
                    for (int m = 0; m < mo; ++m) {
                        const int st = site[q * mo + m];
                        if (st >= 0) p *= s[st];
                    }
                    e += p;
                }
                E_out[r] = e;
                int i = k - 1;                                   /* next k-subset in lexicographic order */
                while (i >= 0 && c[i] == n - k + i) --i;
                if (i < 0) break;
                ++c[i];
                for (int j = i + 1; j < k; ++j) c[j] = c[j - 1] + 1;
            }
        }
    }
    return 0;
}

/* ── single-thread restatement of mchammer's canonical swap step (the reference's CPU sampler benchmark,
 * scripts/bench/bench_stage_mchammer.py:97-120: CanonicalEnsemble.get_swapped_state / _get_property_change /
 * _acceptance_condition / update_occupations) on a binary one-sublattice cell: site a uniform, site b uniform among
 * the sites of the other species (a swap always changes the configuration), the local energy change of the two
 * sites, Metropolis at beta, occupations and per-species site lists updated on acceptance.  Uniforms from Philox
 * counter (t, 0, 0x4D43, 0) under the key (seed).  Returns accepted swaps; E is advanced. */
int64_t oracle_mchammer_swaps(const oracle_ce *m, int8_t *occ, int64_t N, int64_t n_steps, double beta,
                              uint64_t seed, double *E) {
    int64_t *list[2], cnt[2] = {0, 0};
    int64_t *where = (int64_t *)malloc(sizeof(int64_t) * (size_t)N);
    list[0] = (int64_t *)malloc(sizeof(int64_t) * (size_t)N);
    list[1] = (int64_t *)malloc(sizeof(int64_t) * (size_t)N);
    if (!where || !list[0] || !list[1]) { free(where); free(list[0]); free(list[1]); return -1; }
    for (int64_t i = 0; i < N; ++i) { const int s = occ[i] & 1; where[i] = cnt[s]; list[s][cnt[s]++] = i; }
    int64_t acc = 0;
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int64_t t = 0; t < n_steps && cnt[0] > 0 && cnt[1] > 0; ++t) {
        uint32_t ctr[4] = {(uint32_t)t, (uint32_t)((uint64_t)t >> 32), 0x4D43u, 0u}, w[4];
        oracle_philox4x32_10(ctr, key, w);
        const int64_t a = (int64_t)(((uint64_t)w[0] * (uint64_t)N) >> 32);
        const int sa = occ[a] & 1, sb = 1 - sa;
        const int64_t b = list[sb][(int64_t)(((uint64_t)w[1] * (uint64_t)cnt[sb]) >> 32)];
        const double u = ((double)((((uint64_t)w[3] << 32) | w[2]) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        const double dE = oracle_ce_delta(m, occ, a, b, sb, sa);
        if (log(u) < -beta * dE) {
            occ[a] = (int8_t)sb; occ[b] = (int8_t)sa; *E += dE; ++acc;
            const int64_t wa = where[a], wb = where[b];       /* a moves to list sb at b's slot, b to list sa */
            list[sb][wb] = a; where[a] = wb;
            list[sa][wa] = b; where[b] = wa;
        }
    }
    free(where); free(list[0]); free(list[1]);
    return acc;
}
