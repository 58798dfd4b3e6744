// libneuralce_hip.so — gfx950 kernels and C ABI for neuralce's parallel-tempering canonical-swap Monte Carlo.
//
// Kernels (one translation unit):
//   ce_chunk_kernel   n_steps swap-Metropolis steps of one replica per workgroup (persistent over the chunk).
//                     Per step: Philox proposal, gather the two covering windows (int8 occupations, integer
//                     arithmetic on the Smith-normal-form cell coordinate), build per-slot point-function value
//                     tables (old / new) in LDS, evaluate every reverse entry of both sites from the value tables,
//                     reduce with DPP row reductions + LDS across waves, Metropolis test, two-byte write.
//   ce_energy_kernel  full energies (forward rows over every cell), fixed-order two-pass reduction.
//   ce_delta_kernel   swap energy change of explicit proposals (same window path as the chunk kernel).
//   pt_exchange_kernel temperature-swap replica exchange of each job.
// Reference functions replaced: see include/neuralce_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <unordered_map>
#include <unordered_set>
#include <utility>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/neuralce_hip.h"

// ───────────────────────────────────────────── errors ─────────────────────────────────────────────
static thread_local std::string g_err;

static int set_err(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(x)                                                                                    \
    do {                                                                                              \
        hipError_t e_ = (x);                                                                          \
        if (e_ != hipSuccess)                                                                         \
            return set_err(NCE_ERR_HIP, std::string(#x " failed: ") + hipGetErrorString(e_));         \
    } while (0)

extern "C" const char *nce_last_error(void) { return g_err.c_str(); }

// Pointers read from job / model descriptors are generic; the hot ones are cast to the global address space so
// the compiler emits global_load / global_store (flat accesses also count against lgkmcnt, so every LDS wait would
// drain the in-flight window prefetches).
#define NCE_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ inline NCE_GLOBAL T *gptr(T *p) {
    return (NCE_GLOBAL T *)p;
}
template <typename T>
__device__ inline const NCE_GLOBAL T *gptr(const T *p) {
    return (const NCE_GLOBAL T *)p;
}

// ───────────────────────────────────────────── Philox ─────────────────────────────────────────────
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11).
__host__ __device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
}

extern "C" void nce_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    philox4x32_10(c, key[0], key[1]);
    for (int i = 0; i < 4; ++i) out[i] = c[i];
}

// ───────────────────────────────────────────── model ──────────────────────────────────────────────
struct ModelDev {
    int32_t nB, J, Wmax, Wcap, mo, Rt, n_act, max_act_b;
    int32_t d0, d1, d2, n_sl;
    int64_t n_cells, n_sites;
    double factor, zerolet;
    const int32_t *J_b, *W_b;
    const uint2 *ent_slot;      // [nB*J] four u16 slots, padding -> Wcap-1
    const double *ent_coef;     // [nB*J]
    const float *ent_coef_f;    // [nB*J]
    const int4 *win;            // [nB*Wmax] (basis, k0, k1, k2)
    const double *pf;           // [n_sl*4]
    const int32_t *basis_to_sl; // [nB]
    const int32_t *act_b, *act_nb;
    const double *act_cum;
    const int4 *rows;           // [Rt*mo]
    const int32_t *row_bf;      // [Rt*mo]
    const double *row_coef;     // [Rt]
    // division by d1*d2 and d2 of non-negative int32 values (Granlund-Montgomery: q = (hi + ((x - hi) >> 1)) >> sh)
    uint32_t div12_m, div12_s, div2_m, div2_s;
    // binary fast path (one basis site, point functions -1 / +1, window <= 192 slots): entry tables of the
    // speculative kernel, 64 entries per iteration in three classes (FsLo / FsFull / FsTail), per precision
    int32_t fs_ok, fs_nlo, fs_nfull, fs_ntail;
    const unsigned char *fs_tab32, *fs_tab64;
    // full energies: the FP64 layout over one entry per cluster (the entry centred on the cluster's first member)
    int32_t fse_nlo, fse_nfull, fse_ntail;
    const unsigned char *fs_tab_e;
};

// NeuralCE model on the device (kernels in nn_update.hip)
constexpr int NN_MAXL = 8;

struct NNModelDev {
    int32_t d0, d1, d2, nA;
    int64_t n_cells, n_sites;
    uint32_t div12_m, div12_s, div2_m, div2_s;
    int32_t W0, J0, P, n_part, C0, L, act, readout, has_gate;
    int32_t cache_h0;            // 1: the per-replica cache starts with the layer-0 rows h0s [N][C0]
    int32_t cin[NN_MAXL], dd[NN_MAXL], mcat[NN_MAXL];
    int64_t off_self[NN_MAXL], off_shell[NN_MAXL], off_gate[NN_MAXL], off_b[NN_MAXL], off_gb[NN_MAXL];
    int64_t off_wcat[NN_MAXL];   // per layer [cin][mcat]: self | shell_p / n_p | gate_p / n_p, row-major
    const int4 *win;
    const uint2 *ent_slot;       // four u16 slots, padding = 0xFFFF
    const int32_t *ent_col;
    const double *ent_coef;
    const double *ent_q;         // [J0] coefficient / sd of the entry's channel
    const int32_t *shell_count;  // [P]
    const int32_t *part_shell;   // [n_part] shell of each partner
    const int4 *part;            // [n_part]
    int32_t c0_eff;              // layer-0 channels with a non-zero row in the first projection
    const int32_t *c0_ch;        // [c0_eff]
    const double *mu, *sd, *lin, *wts, *head_w;
    const float *mu_f, *sd_f, *lin_f, *wts_f, *head_w_f;
    double head_b, de_scale;
    int32_t ball_cap[NN_MAXL + 1];
    int32_t hash_cap;            // power of two >= 1.5 * ball_cap[L]
    int32_t hash_cap_f;          // the BF16 layers' table (uint16 slots, no hit counts): as large as LDS allows
    // Two flips whose last balls share no site (ib - ia outside ball_L - ball_L): the union's slots, hits and
    // target ranges are then the same for every such pair, so they are tables built once for centre 0.
    // pre_ok = 0: every step builds its balls in the hash table.
    int32_t pre_ok;
    int32_t pre_n[NN_MAXL + 1];  // one flip's ball sizes
    const uint32_t *pre_near;    // [ceil(N / 32)] bit v set: cell vector v lies in ball_L - ball_L
    const int4 *pre_slot;        // [2 pre_n[L]] (flip, offset) of every union slot
    const int32_t *pre_win;      // [2 W0] union slot of every window slot
    const int32_t *pre_list;     // per layer: hits (source slot << 8 | shell), sorted by target
    const int32_t *pre_end;      // per layer: end of each target's range in pre_list
    int64_t pre_list_off[NN_MAXL], pre_end_off[NN_MAXL];
    // per precision: which per-step arrays live in LDS (else in the scratch): 0 none, 1 the layer-0 change of
    // ball 0, 2 and the sorted partner hits, 3 and the projections
    int32_t lds_mode32, lds_mode64, lds_mode16;
    int32_t fp32_in_lds;         // the whole FP32 step fits in LDS (placement 3 before any cap from the environment)
    int64_t pr_elems, hit_elems, w_elems;   // largest projection block / hit list / used projection matrix
    int32_t mfma_ok;             // every layer's projection width is a multiple of 16 (FP32 MFMA tiles)
    // BF16 layers (gathered shell sums times the stacked matrix on MFMA): every width D a multiple of 16,
    // D (1 + gate) <= 128 and every input width <= 64.  Per layer the input width padded to 32 (cpad), the stacked
    // matrix's k extent
    // cpad (1 + 2P) (self, then each shell's sum as a bf16 high and low part) and its LDS row stride (+8).
    int32_t fused_ok, xb_stride, xb_zrow;   // xb_zrow: the all-zero row of either input-row buffer
    int32_t cpad[NN_MAXL], kf[NN_MAXL];
    int64_t wf_bytes;            // largest stacked matrix in LDS (bf16)
    const uint16_t *wf_img;      // every layer's stacked matrix as bf16 bits, in its LDS layout (row stride kf + 8)
    int64_t wf_off[NN_MAXL];
    // BF16 layers: per target a row of nbr_w = P * nbr_ps uint16 source slots, shell p's partners in columns
    // p * nbr_ps + (index within the shell), the zero row xb_zrow where the partner is outside ball_{l-1}.
    // Fixed balls: one table per layer in device memory (pre_nbr); hash steps build one for every layer in the
    // scratch (a layer takes the sources below its ball's size).
    int32_t nbr_w, nbr_ps;
    const int32_t *part_col;     // [n_part] column of each partner in a neighbour-table row
    const uint16_t *pre_nbr;
    int64_t pre_nbr_off[NN_MAXL], soff_nbr;
    int64_t soff_pub;            // split replicas: slots published by the part that builds a hash step's balls
    int64_t xb_off1;             // the BF16 layers' two input-row buffers (bf16 [rows][xb_stride]) start at the
                                 // projection block; the second this many bf16 elements after the first
    int64_t off_pre[NN_MAXL], off_gpre[NN_MAXL], cache_elems;
    int64_t soff_dh0, soff_pr, soff_tgt, soff_list, soff_npre[NN_MAXL], soff_ngpre[NN_MAXL], soff_dh[2], scratch_elems;
};

struct NNModelHost {
    NNModelDev m;
};

// Partial sums of nce_ce_energy_noalloc: file-scope device memory, one copy per GPU that loads the library, held
// until the process ends (8 MiB).  In host code `scratch` only names the symbol: its device address comes from
// hipGetSymbolAddress at context creation (kept in nce_ctx::noalloc_partial) and is what kernels receive.
static __device__ __align__(256) double scratch[NCE_NOALLOC_PARTIALS];

struct nce_ctx {
    int device = 0;
    double *noalloc_partial = nullptr;   // this device's address of `scratch`
    int n_cu = 0;
    size_t max_lds = 0;
    std::vector<ModelDev> models;
    std::vector<size_t> lds_chunk64, lds_chunk32, lds_energy;
    ModelDev *d_models = nullptr;
    int cap_models = 0;
    std::vector<void *> allocs;
    struct Stage {
        void *host = nullptr, *dev = nullptr;
        size_t cap = 0;
        hipEvent_t ev = nullptr;
        bool pending = false;
    };
    // Staged uploads of job tables: one ring per stream.  A slot's next copy is ordered after the kernels that read
    // it by the stream itself; the slot's event (after its copy) guards the pinned host buffer.  Rings shared by
    // streams would let a copy on one stream overwrite a table that a long launch on another still reads.
    struct StageRing {
        hipStream_t stream = nullptr;
        Stage stage[4];
        int next = 0;
    };
    std::vector<StageRing *> rings;
    // Per-block partial sums of the full-energy kernels, one buffer per stream (like the staging rings: a buffer
    // shared by streams could be rewritten on one stream while a launch on another still reads it).
    struct Partial {
        hipStream_t stream = nullptr;
        double *p = nullptr;
        size_t cap = 0;
    };
    std::vector<Partial> partials;
    // layer-0 rows of NeuralCE full recomputes for models whose caches do not keep them, one buffer per stream
    std::vector<Partial> h0_rows;
    // barrier counters and energy partials of NeuralCE replicas split over several workgroups, one buffer per stream
    std::vector<Partial> split_mem;
    std::vector<NNModelDev> nn_models;
    void *d_nn_models = nullptr;
    int cap_nn_models = 0;
};

__host__ __device__ inline size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

// LDS layout of the chunk / delta kernels (identical carve on host and device)
struct ChunkLds {
    size_t slot, coef, win, valA, valB, red, total;
};

template <typename Real>
__host__ __device__ inline ChunkLds chunk_lds_layout(int nB, int J, int Wmax, int Wcap, int waves) {
    ChunkLds L;
    size_t o = 0;
    L.slot = o; o = align16(o + (size_t)nB * J * sizeof(uint2));
    L.coef = o; o = align16(o + (size_t)nB * J * sizeof(Real));
    L.win = o;  o = align16(o + (size_t)nB * Wmax * sizeof(int4));
    L.valA = o; o = align16(o + (size_t)Wcap * 2 * sizeof(Real));
    L.valB = o; o = align16(o + (size_t)Wcap * 2 * sizeof(Real));
    L.red = o;  o = align16(o + (size_t)waves * sizeof(double) * 2);
    L.total = o;
    return L;
}

// ───────────────────────────────────────────── reductions ─────────────────────────────────────────
// Row (16-lane) butterfly with DPP: quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror, row_mirror.
template <int CTRL>
__device__ inline double dpp_f64(double v) {
    const int64_t b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffff), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((int64_t)hi << 32) | (uint32_t)lo);
}
template <int CTRL>
__device__ inline float dpp_f32(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
__device__ inline double readlane_f64(double v, int lane) {
    const int64_t b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((int64_t)hi << 32) | (uint32_t)lo);
}

// workgroup barrier that orders LDS only: global loads stay in flight across it (a plain __syncthreads waits for
// every outstanding load and store of the wave first)
__device__ inline void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// DPP copy for scans: lanes whose source lies outside the row, or whose row / bank the masks leave out, read 0
template <int CTRL, int ROWS, int BANKS>
__device__ inline double dpp_f64_z(double v) {
    const int64_t b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffff), CTRL, ROWS, BANKS, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWS, BANKS, true);
    return __longlong_as_double(((int64_t)hi << 32) | (uint32_t)lo);
}
// inclusive prefix sum over the 64 lanes: row_shr 1, 2, 4, 8 within 16-lane rows, then row_bcast:15 into rows 1
// and 3 and row_bcast:31 into rows 2 and 3; fixed order (deterministic)
__device__ inline double wave_prefix_sum(double v) {
    v += dpp_f64_z<0x111, 0xF, 0xF>(v);
    v += dpp_f64_z<0x112, 0xF, 0xF>(v);
    v += dpp_f64_z<0x114, 0xF, 0xF>(v);
    v += dpp_f64_z<0x118, 0xF, 0xF>(v);
    v += dpp_f64_z<0x142, 0xA, 0xF>(v);
    v += dpp_f64_z<0x143, 0xC, 0xF>(v);
    return v;
}
__device__ inline double wave_min(double v) {
    v = fmin(v, dpp_f64<0xB1>(v));
    v = fmin(v, dpp_f64<0x4E>(v));
    v = fmin(v, dpp_f64<0x141>(v));
    v = fmin(v, dpp_f64<0x140>(v));
    return fmin(fmin(readlane_f64(v, 0), readlane_f64(v, 16)), fmin(readlane_f64(v, 32), readlane_f64(v, 48)));
}

// wave sum, result uniform in every lane; fixed order (deterministic)
__device__ inline double wave_sum(double v) {
    v += dpp_f64<0xB1>(v);
    v += dpp_f64<0x4E>(v);
    v += dpp_f64<0x141>(v);
    v += dpp_f64<0x140>(v);
    return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}
__device__ inline float wave_sum(float v) {
    v += dpp_f32<0xB1>(v);
    v += dpp_f32<0x4E>(v);
    v += dpp_f32<0x141>(v);
    v += dpp_f32<0x140>(v);
    const int i = __float_as_int(v);
    return (__int_as_float(__builtin_amdgcn_readlane(i, 0)) + __int_as_float(__builtin_amdgcn_readlane(i, 16))) +
           (__int_as_float(__builtin_amdgcn_readlane(i, 32)) + __int_as_float(__builtin_amdgcn_readlane(i, 48)));
}

template <typename Real, int WAVES>
__device__ inline Real block_sum(Real v, Real *red) {
    v = wave_sum(v);
    if constexpr (WAVES == 1) {
        return v;
    } else {
        const int w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) red[w] = v;
        __syncthreads();
        Real s = red[0];
#pragma unroll
        for (int i = 1; i < WAVES; ++i) s += red[i];
        return s;
    }
}

// ───────────────────────────────────────────── swap energy change ─────────────────────────────────
__device__ inline uint32_t udiv_magic(uint32_t x, uint32_t m, uint32_t sh, uint32_t d) {
    if (m == 0) return d == 1 ? x : x >> sh;                          // d == 1
    const uint32_t hi = __umulhi(x, m);
    return (hi + ((x - hi) >> 1)) >> sh;
}

struct Coord {
    int32_t b, k0, k1, k2;
};

__device__ inline Coord site_coord(const ModelDev &m, int64_t i) {
    Coord c;
    c.b = (int32_t)(i % m.nB);
    const int64_t cell = i / m.nB;
    const int64_t d12 = (int64_t)m.d1 * m.d2;
    c.k0 = (int32_t)(cell / d12);
    const int64_t rem = cell - (int64_t)c.k0 * d12;
    c.k1 = (int32_t)(rem / m.d2);
    c.k2 = (int32_t)(rem - (int64_t)c.k1 * m.d2);
    return c;
}

__device__ inline int64_t window_site(const ModelDev &m, const Coord &c, int4 e) {
    int32_t k0 = c.k0 + e.y, k1 = c.k1 + e.z, k2 = c.k2 + e.w;
    k0 -= (k0 >= m.d0) ? m.d0 : 0;
    k1 -= (k1 >= m.d1) ? m.d1 : 0;
    k2 -= (k2 >= m.d2) ? m.d2 : 0;
    return (((int64_t)k0 * m.d1 + k1) * m.d2 + k2) * m.nB + e.x;
}

template <typename Real>
struct Real2 {
    Real o, n;
};

// Fill the value tables of both windows for the swap (ia: sa -> sb, ib: sb -> sa).
//   window A (entries of ia): slot value (pf(old), pf(new)) with both overrides applied by site index;
//   window B (entries of ib): same, except slots that are the site ia get (0, 0), so every entry of ib that also
//   contains ia contributes nothing (the reference skips those entries, ce_gpu_implicit.py:189-191).
template <typename Real, int BS>
__device__ inline void fill_values(const ModelDev &m, const int4 *s_win, Real2<Real> *valA, Real2<Real> *valB,
                                   const NCE_GLOBAL int8_t *__restrict__ occ, int64_t ia, int64_t ib, int sa, int sb) {
    const Coord ca = site_coord(m, ia), cb = site_coord(m, ib);
    const int Wa = m.W_b[ca.b], Wb = m.W_b[cb.b];
    for (int w = threadIdx.x; w < Wa + Wb; w += BS) {
        const bool isA = w < Wa;
        const int slot = isA ? w : w - Wa;
        const Coord &c = isA ? ca : cb;
        const int4 e = s_win[c.b * m.Wmax + slot];
        const int64_t s = window_site(m, c, e);
        const int sp = occ[s];
        int spn = sp;
        if (s == ia) spn = sb;
        if (s == ib) spn = sa;
        const double *pf = m.pf + m.basis_to_sl[e.x] * 4 + 2;       // bf = 1
        Real2<Real> v;
        v.o = (Real)pf[sp];
        v.n = (Real)pf[spn];
        if (!isA && s == ia) { v.o = (Real)0; v.n = (Real)0; }
        (isA ? valA : valB)[slot] = v;
    }
}

template <typename Real>
__device__ inline Real entry_delta(const Real2<Real> *val, uint2 e) {
    const Real2<Real> a = val[e.x & 0xffff], b = val[e.x >> 16], c = val[e.y & 0xffff], d = val[e.y >> 16];
    return (a.n * b.n) * (c.n * d.n) - (a.o * b.o) * (c.o * d.o);
}

// the same for binary point functions held as int8 (-1 / +1, 0 for the excluded slots): exact in integers
__device__ inline int entry_delta_bin(const char2 *val, uint2 e) {
    const char2 a = val[e.x & 0xffff], b = val[e.x >> 16], c = val[e.y & 0xffff], d = val[e.y >> 16];
    return (a.y * b.y) * (c.y * d.y) - (a.x * b.x) * (c.x * d.x);
}

// Partial sum of coef * (prod_new - prod_old) over both sites' entries for this thread.
template <typename Real, int BS>
__device__ inline Real entries_partial(const ModelDev &m, const uint2 *s_slot, const Real *s_coef,
                                       const Real2<Real> *valA, const Real2<Real> *valB, int ba, int bb) {
    Real acc = (Real)0;
    if (ba == bb) {
        const int Jb = m.J_b[ba];
        const uint2 *sl = s_slot + (size_t)ba * m.J;
        const Real *cf = s_coef + (size_t)ba * m.J;
        for (int j = threadIdx.x; j < Jb; j += BS) {
            const uint2 e = sl[j];
            acc += cf[j] * (entry_delta(valA, e) + entry_delta(valB, e));
        }
    } else {
        for (int pass = 0; pass < 2; ++pass) {
            const int b = pass ? bb : ba;
            const Real2<Real> *val = pass ? valB : valA;
            const int Jb = m.J_b[b];
            const uint2 *sl = s_slot + (size_t)b * m.J;
            const Real *cf = s_coef + (size_t)b * m.J;
            for (int j = threadIdx.x; j < Jb; j += BS) acc += cf[j] * entry_delta(val, sl[j]);
        }
    }
    return acc;
}

template <typename Real, int BS>
__device__ inline void load_model_lds(const ModelDev &m, uint2 *s_slot, Real *s_coef, int4 *s_win,
                                      Real2<Real> *valA, Real2<Real> *valB) {
    const int nJ = m.nB * m.J;
    for (int i = threadIdx.x; i < nJ; i += BS) {
        s_slot[i] = m.ent_slot[i];
        if constexpr (sizeof(Real) == 8) s_coef[i] = (Real)m.ent_coef[i];
        else s_coef[i] = (Real)m.ent_coef_f[i];
    }
    for (int i = threadIdx.x; i < m.nB * m.Wmax; i += BS) s_win[i] = m.win[i];
    if (threadIdx.x == 0) {                         // the constant slot: point-function value 1 before and after
        valA[m.Wcap - 1].o = valA[m.Wcap - 1].n = (Real)1;
        valB[m.Wcap - 1].o = valB[m.Wcap - 1].n = (Real)1;
    }
}

// ───────────────────────────────────────────── proposal ───────────────────────────────────────────
struct Proposal {
    int64_t ia, ib;
    double u;
};

__device__ inline uint64_t umulhi64(uint64_t a, uint64_t b) { return __umul64hi(a, b); }

// Stream layout shared with the oracle: x = philox({step, replica, 0}), y = philox({step, replica, 1}).
__device__ inline Proposal draw_proposal(const ModelDev &m, uint64_t seed, uint64_t step, uint32_t rep) {
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    uint32_t x[4] = {(uint32_t)step, (uint32_t)(step >> 32), rep, 0u};
    uint32_t y[4] = {(uint32_t)step, (uint32_t)(step >> 32), rep, 1u};
    philox4x32_10(x, k0, k1);
    philox4x32_10(y, k0, k1);
    const uint64_t ca = umulhi64(((uint64_t)x[1] << 32) | x[0], (uint64_t)m.n_cells);
    const uint64_t cb = umulhi64(((uint64_t)x[3] << 32) | x[2], (uint64_t)m.n_cells);
    int sl = 0;
    if (m.n_act > 1) {
        const double us = (double)y[2] * (1.0 / 4294967296.0);
        while (sl < m.n_act - 1 && !(us < m.act_cum[sl])) ++sl;
    }
    const uint32_t nb = (uint32_t)m.act_nb[sl];
    const int ba = m.act_b[sl * m.max_act_b + (int)(((y[3] & 0xffffu) * nb) >> 16)];
    const int bb = m.act_b[sl * m.max_act_b + (int)(((y[3] >> 16) * nb) >> 16)];
    Proposal p;
    p.ia = (int64_t)ca * m.nB + ba;
    p.ib = (int64_t)cb * m.nB + bb;
    p.u = ((double)((((uint64_t)y[1] << 32) | y[0]) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    return p;
}

// Same stream as draw_proposal for one active sublattice with one basis site (site index = cell index): the
// sublattice and basis draws of y2, y3 are trivial there, so only n_cells is needed (no table reads).
__device__ inline Proposal draw_proposal_1b(uint64_t seed, uint64_t step, uint32_t rep, uint64_t n_cells) {
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    uint32_t x[4] = {(uint32_t)step, (uint32_t)(step >> 32), rep, 0u};
    uint32_t y[4] = {(uint32_t)step, (uint32_t)(step >> 32), rep, 1u};
    philox4x32_10(x, k0, k1);
    philox4x32_10(y, k0, k1);
    Proposal p;
    p.ia = (int64_t)umulhi64(((uint64_t)x[1] << 32) | x[0], n_cells);
    p.ib = (int64_t)umulhi64(((uint64_t)x[3] << 32) | x[2], n_cells);
    p.u = ((double)((((uint64_t)y[1] << 32) | y[0]) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    return p;
}

// ───────────────────────────────────────────── chunk kernel ───────────────────────────────────────
template <typename Real, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void ce_chunk_kernel(const ModelDev *__restrict__ models,
                                                              const nce_pt_job *__restrict__ jobs,
                                                              const int2 *__restrict__ items, int64_t n_steps) {
    constexpr int BS = 64 * WAVES;
    extern __shared__ __align__(16) unsigned char smem[];
    const int2 it = items[blockIdx.x];
    const nce_pt_job job = jobs[it.x];
    const ModelDev m = models[job.model];
    const int r = it.y;
    const ChunkLds L = chunk_lds_layout<Real>(m.nB, m.J, m.Wmax, m.Wcap, WAVES);
    uint2 *s_slot = (uint2 *)(smem + L.slot);
    Real *s_coef = (Real *)(smem + L.coef);
    int4 *s_win = (int4 *)(smem + L.win);
    Real2<Real> *valA = (Real2<Real> *)(smem + L.valA);
    Real2<Real> *valB = (Real2<Real> *)(smem + L.valB);
    Real *red = (Real *)(smem + L.red);
    load_model_lds<Real, BS>(m, s_slot, s_coef, s_win, valA, valB);
    __syncthreads();

    NCE_GLOBAL int8_t *occ = gptr(job.occ) + (int64_t)r * job.n_sites;
    double E = job.energy[r];
    const double beta = job.beta[r];
    const float beta_f = (float)beta;
    const int slot = job.slot_of_rep[r];
    const uint64_t step0 = (uint64_t)job.step[r];
    double emin = job.e_min[r];
    const double eref = job.accumulate ? job.e_ref[r] : 0.0;
    double s1 = 0.0, s2 = 0.0;
    int64_t nacc = 0;
    const int64_t inj_base = (int64_t)r * n_steps;

    for (int64_t t = 0; t < n_steps; ++t) {
        Proposal p;
        if (job.inj_ia) {
            p.ia = job.inj_ia[inj_base + t];
            p.ib = job.inj_ib[inj_base + t];
            p.u = job.inj_u[inj_base + t];
        } else {
            p = draw_proposal(m, job.seed, step0 + (uint64_t)t, (uint32_t)r);
        }
        const int sa = occ[p.ia], sb = occ[p.ib];
        if (sa != sb) {
            fill_values<Real, BS>(m, s_win, valA, valB, occ, p.ia, p.ib, sa, sb);
            __syncthreads();
            const int ba = (int)(p.ia % m.nB), bb = (int)(p.ib % m.nB);
            Real part = entries_partial<Real, BS>(m, s_slot, s_coef, valA, valB, ba, bb);
            const Real sum = block_sum<Real, WAVES>(part, red);
            bool accept;
            double dE;
            if constexpr (sizeof(Real) == 8) {
                dE = (double)sum * m.factor;
                accept = log(p.u) < -beta * dE;
            } else {
                const float dEf = (float)sum * (float)m.factor;
                dE = (double)dEf;
                accept = logf((float)p.u) < -(beta_f * dEf);
            }
            if (accept) {
                if (threadIdx.x == 0) {
                    occ[p.ia] = (int8_t)sb;
                    occ[p.ib] = (int8_t)sa;
                }
                E += dE;
                ++nacc;
            }
        }
        emin = fmin(emin, E);
        if (job.accumulate) {
            const double x = E - eref;
            s1 += x;
            s2 += x * x;
        }
        if (job.traj && ((t + 1) % job.sample_interval) == 0 && threadIdx.x == 0)
            job.traj[((t + 1) / job.sample_interval - 1) * job.n_replicas + slot] = E;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        job.energy[r] = E;
        job.e_min[r] = emin;
        job.n_accept[r] += nacc;
        job.step[r] = (int64_t)(step0 + (uint64_t)n_steps);
        if (job.accumulate) {
            job.s1[r] += s1;
            job.s2[r] += s2;
        }
    }
}

// ───────────────────────────────────────────── binary fast chunk kernel ──────────────────────────
// Binary point functions (-1 / +1): a cluster's product is (-1)^(number of members whose species is 0).  Away from
// each other and from their own periodic images, swapping two different species flips exactly the centre of each
// window, so dE = -2 * factor * (h_a + h_b) with h_x = sum over the entries e of x of coef_e * prod_e.  Splitting off
// the centre, prod_e = (-1)^[centre is species 0] * (-1)^parity(bits_x & mask_e), mask_e = the other members' window
// bits; and since the two centres hold different species, h_a + h_b = (-1)^[s_a = 0] * (S_a - S_b) with
// S_x = sum_e coef_e * (-1)^parity(bits_x & mask_e).  Bit 0 of window word 0 (the centre) is set to 1 in bits_x
// and to [coef_e < 0] in mask_e, so the table stores |coef_e| and a term is one AND + popcount + shift-or of the
// parity into the sign bit.  Steps whose windows contain the other site or an image of their own centre take the
// value-table path.
template <typename Real>
struct FsLo;                                   // every member in window bits 0..31
template <>
struct FsLo<float> {
    uint32_t m, c;
};
template <>
struct FsLo<double> {
    uint32_t m, pad;
    uint64_t c;
};
struct FsFull {                                // members in window word 0
    uint64_t m, c;                             // c: |coef| bits (float in the low half)
};
struct FsTail {                                // a member in window word `word` (1 or 2)
    uint64_t m0, mk, c;
    uint32_t word, pad;
};
template <typename Real>
__host__ __device__ inline size_t fs_bytes(int nlo, int nfull, int ntail) {
    return (size_t)64 * ((size_t)nlo * sizeof(FsLo<Real>) + (size_t)nfull * sizeof(FsFull) +
                         (size_t)ntail * sizeof(FsTail));
}

// (a & m) ^ t in one v_bitop3 (truth table over S0 = 0xF0, S1 = 0xCC, S2 = 0xAA)
__device__ inline uint32_t and_xor(uint32_t a, uint32_t m, uint32_t t) {
    return __builtin_amdgcn_bitop3_b32(a, m, t, 0x6A);
}
// |c| (sign bit clear) with the low bit of n as its sign
__device__ inline float fs_term(uint32_t c, uint32_t n) { return __int_as_float((int)((n << 31) | c)); }
__device__ inline double fs_term(uint64_t c, uint32_t n) {
    return __longlong_as_double((long long)(((uint64_t)(n << 31) << 32) | c));
}

// Speculative parallel steps: a group is the run of steps holding the next W different-species swaps of one
// replica; the wave of rank k evaluates the group's k-th swap from the current occupations, so the window loads of
// all W swaps are in flight together.  The decisions are then taken in step order (Metropolis, running total,
// statistics); a swap whose windows contain a site changed by an earlier accepted swap of the same group is
// re-evaluated on patched species before its decision, so the result is exactly the sequential chain.
struct SpecRec {
    double dE;
    double lu;                       // log of the step's Metropolis uniform, in the precision of the decision
    int sa, sb, valid, conflict;
};

template <typename Real>
struct SpecLds {
    size_t slot, coef, val, prop, rec, commit, fs, pp, meta, total;
};

constexpr int SPEC_KP = 64;            // proposals scanned per group (one lane of wave 0 each); W of them kept

struct SpecProp {
    int ia, ib;
    double u;
};

template <typename Real>
__host__ __device__ inline SpecLds<Real> spec_lds_layout(int J, int Wcap, int waves, int nlo, int nfull, int ntail) {
    SpecLds<Real> L;
    size_t o = 0;
    L.slot = o;   o = align16(o + (size_t)J * sizeof(uint2));
    L.coef = o;   o = align16(o + (size_t)J * sizeof(Real));
    L.val = o;    o = align16(o + (size_t)waves * 2 * Wcap * sizeof(char2));
    L.prop = o;   o = align16(o + (size_t)waves * sizeof(int2));
    L.rec = o;    o = align16(o + (size_t)waves * sizeof(SpecRec));
    L.commit = o; o = align16(o + (size_t)waves * sizeof(int4));
    L.fs = o;     o = align16(o + fs_bytes<Real>(nlo, nfull, ntail));
    L.pp = o;     o = align16(o + (size_t)waves * sizeof(SpecProp));
    L.meta = o;   o = align16(o + 2 * sizeof(unsigned long long));
    L.total = o;
    return L;
}

// Step rank of wave w within a group.  Wave w runs on SIMD w % 4 and the rank-k step scans the k earlier steps for
// conflicts, so the ranks are spread such that every SIMD's ranks add up to the same total.
template <int W>
__device__ inline int spec_rank(int w) {
    const int s = w & 3, slot = w >> 2;
    if constexpr (W == 4) return w;
    else if constexpr (W == 8) return slot == 0 ? s : 7 - s;
    else return slot == 0 ? s : slot == 1 ? 7 - s : slot == 2 ? 8 + s : 15 - s;
}

template <typename Real, int W, bool INJ>
__global__ __launch_bounds__(64 * W) void ce_chunk_spec_kernel(const ModelDev *__restrict__ models,
                                                               const nce_pt_job *__restrict__ jobs,
                                                               const int2 *__restrict__ items, int64_t n_steps) {
    constexpr int BS = 64 * W;
    constexpr int NS = 3;                              // window slots per lane: lane, lane + 64, lane + 128
    extern __shared__ __align__(16) unsigned char smem[];
    const int2 it = items[blockIdx.x];
    const nce_pt_job job = jobs[it.x];
    const ModelDev &mg = models[job.model];
    const int r = it.y;
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int kw = spec_rank<W>(w);                   // this wave's step in every group
    const int J = mg.J, Wcap = mg.Wcap, Wn = mg.W_b[0], Jb0 = mg.J_b[0];
    const int nlo = mg.fs_nlo, nfull = mg.fs_nfull, ntail = mg.fs_ntail;
    const SpecLds<Real> L = spec_lds_layout<Real>(J, Wcap, W, nlo, nfull, ntail);
    uint2 *s_slot = (uint2 *)(smem + L.slot);
    Real *s_coef = (Real *)(smem + L.coef);
    char2 *valA = (char2 *)(smem + L.val) + (size_t)w * 2 * Wcap;   // (old, new) point-function values
    char2 *valB = valA + Wcap;
    int2 *s_prop = (int2 *)(smem + L.prop);
    SpecRec *s_rec = (SpecRec *)(smem + L.rec);
    int4 *s_commit = (int4 *)(smem + L.commit);
    SpecProp *s_pp = (SpecProp *)(smem + L.pp);
    unsigned long long *s_meta = (unsigned long long *)(smem + L.meta);
    const FsLo<Real> *s_lo = (const FsLo<Real> *)(smem + L.fs);
    const FsFull *s_full = (const FsFull *)(s_lo + 64 * nlo);
    const FsTail *s_tail = (const FsTail *)(s_full + 64 * nfull);
    for (int i = tid; i < J; i += BS) {
        s_slot[i] = mg.ent_slot[i];
        if constexpr (sizeof(Real) == 8) s_coef[i] = mg.ent_coef[i];
        else s_coef[i] = mg.ent_coef_f[i];
    }
    {
        const uint4 *src = (const uint4 *)(sizeof(Real) == 8 ? mg.fs_tab64 : mg.fs_tab32);
        uint4 *dst = (uint4 *)(smem + L.fs);
        const int n16 = (int)(fs_bytes<Real>(nlo, nfull, ntail) / 16);
        for (int i = tid; i < n16; i += BS) dst[i] = src[i];
    }
    if (lane == 0) {
        valA[Wcap - 1] = make_char2(1, 1);
        valB[Wcap - 1] = make_char2(1, 1);
    }
    int4 wofs[NS];
    bool wok[NS];
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        const int slot = lane + 64 * j;
        wok[j] = slot < Wn;
        wofs[j] = make_int4(0, 0, 0, 0);
        if (wok[j]) {
            const NCE_GLOBAL int *wp = (const NCE_GLOBAL int *)gptr(mg.win) + 4 * slot;
            wofs[j] = make_int4(wp[0], wp[1], wp[2], wp[3]);
        }
    }
    const int64_t n_cells = mg.n_cells;
    const int d0 = mg.d0, d1 = mg.d1, d2 = mg.d2;
    const uint32_t m12 = mg.div12_m, s12 = mg.div12_s, m2 = mg.div2_m, sh2 = mg.div2_s;
    // window offsets as signed representatives in (-d/2, d/2] and their box: the window of a site at least the box
    // away from every boundary does not wrap, and its slot addresses are the site plus lin[j]
    int lin[NS];
    int blo0 = 0, bhi0 = 0, blo1 = 0, bhi1 = 0, blo2 = 0, bhi2 = 0;
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        lin[j] = 0;
        if (wok[j]) {
            const int o0 = wofs[j].y > d0 / 2 ? wofs[j].y - d0 : wofs[j].y;
            const int o1 = wofs[j].z > d1 / 2 ? wofs[j].z - d1 : wofs[j].z;
            const int o2 = wofs[j].w > d2 / 2 ? wofs[j].w - d2 : wofs[j].w;
            lin[j] = (o0 * d1 + o1) * d2 + o2;
            blo0 = min(blo0, o0); bhi0 = max(bhi0, o0);
            blo1 = min(blo1, o1); bhi1 = max(bhi1, o1);
            blo2 = min(blo2, o2); bhi2 = max(bhi2, o2);
        }
    }
    for (int o = 1; o < 64; o <<= 1) {
        blo0 = min(blo0, __shfl_xor(blo0, o, 64)); bhi0 = max(bhi0, __shfl_xor(bhi0, o, 64));
        blo1 = min(blo1, __shfl_xor(blo1, o, 64)); bhi1 = max(bhi1, __shfl_xor(bhi1, o, 64));
        blo2 = min(blo2, __shfl_xor(blo2, o, 64)); bhi2 = max(bhi2, __shfl_xor(bhi2, o, 64));
    }
    blo0 = __builtin_amdgcn_readfirstlane(blo0); bhi0 = __builtin_amdgcn_readfirstlane(bhi0);
    blo1 = __builtin_amdgcn_readfirstlane(blo1); bhi1 = __builtin_amdgcn_readfirstlane(bhi1);
    blo2 = __builtin_amdgcn_readfirstlane(blo2); bhi2 = __builtin_amdgcn_readfirstlane(bhi2);
    // a window holding an image of its own centre (a slot other than 0 at offset 0 modulo the cell): small cells only
    bool selfimg = false;
#pragma unroll
    for (int j = 0; j < NS; ++j)
        selfimg |= wok[j] && (lane + 64 * j) > 0 && wofs[j].y == 0 && wofs[j].z == 0 && wofs[j].w == 0;
    const bool self_image = __ballot(selfimg) != 0ull;
    auto interior = [&](int c0, int c1, int c2) __attribute__((always_inline)) -> bool {
        return c0 + blo0 >= 0 && c0 + bhi0 < d0 && c1 + blo1 >= 0 && c1 + bhi1 < d1 && c2 + blo2 >= 0 &&
               c2 + bhi2 < d2;
    };
    const uint32_t d12 = (uint32_t)(d1 * d2);
    auto coords = [&](int x, int &c0, int &c1, int &c2) __attribute__((always_inline)) {
        c0 = (int)udiv_magic((uint32_t)x, m12, s12, d12);
        const int rem = x - c0 * (int)d12;
        c1 = (int)udiv_magic((uint32_t)rem, m2, sh2, (uint32_t)d2);
        c2 = rem - c1 * d2;
    };
    // site y (coordinates y0..y2) inside the offset box of the window of a site with coordinates c0..c2: necessary
    // for y to be a slot of that window
    auto in_box = [&](int y0, int y1, int y2, int c0, int c1, int c2) __attribute__((always_inline)) -> bool {
        int e0 = y0 - c0, e1 = y1 - c1, e2 = y2 - c2;
        e0 += e0 < 0 ? d0 : 0; e1 += e1 < 0 ? d1 : 0; e2 += e2 < 0 ? d2 : 0;
        return (e0 <= bhi0 || e0 >= d0 + blo0) && (e1 <= bhi1 || e1 >= d1 + blo1) && (e2 <= bhi2 || e2 >= d2 + blo2);
    };
    const double factor = mg.factor;

    NCE_GLOBAL int8_t *occ = gptr(job.occ) + (int64_t)r * job.n_sites;
    const NCE_GLOBAL int64_t *inj_ia = gptr(job.inj_ia), *inj_ib = gptr(job.inj_ib);
    const NCE_GLOBAL double *inj_u = gptr(job.inj_u);
    NCE_GLOBAL double *traj = gptr(job.traj);
    double E = job.energy[r];
    const double beta = job.beta[r];
    const float beta_f = (float)beta;
    const int slot_r = job.slot_of_rep[r];
    const int64_t step0_v = job.step[r];
    const uint64_t step0 = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(step0_v >> 32)) << 32) |
                           (uint32_t)__builtin_amdgcn_readfirstlane((int)step0_v);
    const uint64_t seed = job.seed;
    double emin = job.e_min[r];
    const double eref = job.accumulate ? job.e_ref[r] : 0.0;
    double s1 = 0.0, s2 = 0.0;
    int64_t nacc = 0;
    const int64_t inj_base = (int64_t)r * n_steps;
    const bool accumulate = job.accumulate != 0;
    const int64_t sample_interval = job.sample_interval, n_rep = job.n_replicas;
    int64_t next_sample = sample_interval, n_sampled = 0;

    // evaluation of this wave's step from (patched) species in registers
    auto evaluate = [&](int ia, int ib, const int (&sA)[NS], const int (&sB)[NS], const int (&pA)[NS],
                        const int (&pB)[NS], int sa, int sb, bool near) __attribute__((always_inline)) -> double {
        Real sum;
        if (!near) {
            unsigned long long WA[3], WB[3];
#pragma unroll
            for (int j = 0; j < NS; ++j) {
                WA[j] = __ballot(wok[j] && pA[j] == 0);
                WB[j] = __ballot(wok[j] && pB[j] == 0);
            }
            WA[0] |= 1ull;                             // the sign bit of every term (see above)
            WB[0] |= 1ull;
            using CT = typename std::conditional<sizeof(Real) == 8, uint64_t, uint32_t>::type;
            Real accA = (Real)0, accB = (Real)0;
            const uint32_t a32 = (uint32_t)WA[0], b32 = (uint32_t)WB[0];
            // uniform trip counts: the loop control stays on the scalar unit
            const FsLo<Real> *plo = s_lo + lane;
#pragma unroll 4
            for (int it = 0; it < nlo; ++it) {
                const FsLo<Real> e = plo[64 * it];
                accA += fs_term((CT)e.c, (uint32_t)__builtin_popcount(a32 & e.m));
                accB += fs_term((CT)e.c, (uint32_t)__builtin_popcount(b32 & e.m));
            }
            const uint32_t ahi = (uint32_t)(WA[0] >> 32), bhi = (uint32_t)(WB[0] >> 32);
            const FsFull *pfu = s_full + lane;
#pragma unroll 4
            for (int it = 0; it < nfull; ++it) {
                const FsFull e = pfu[64 * it];
                // parity of a 64-bit AND = parity of its two words XOR-folded (one popcount)
                const uint32_t mlo = (uint32_t)e.m, mhi = (uint32_t)(e.m >> 32);
                accA += fs_term((CT)e.c, (uint32_t)__builtin_popcount(and_xor(a32, mlo, ahi & mhi)));
                accB += fs_term((CT)e.c, (uint32_t)__builtin_popcount(and_xor(b32, mlo, bhi & mhi)));
            }
            for (int it = 0; it < ntail; ++it) {
                const FsTail e = s_tail[64 * it + lane];
                const unsigned long long a1 = e.word == 1 ? WA[1] : WA[2], b1 = e.word == 1 ? WB[1] : WB[2];
                const uint32_t m0l = (uint32_t)e.m0, m0h = (uint32_t)(e.m0 >> 32);
                const uint32_t mkl = (uint32_t)e.mk, mkh = (uint32_t)(e.mk >> 32);
                const uint32_t fa = and_xor(a32, m0l, and_xor(ahi, m0h, and_xor((uint32_t)a1, mkl,
                                                                                (uint32_t)(a1 >> 32) & mkh)));
                const uint32_t fb = and_xor(b32, m0l, and_xor(bhi, m0h, and_xor((uint32_t)b1, mkl,
                                                                                (uint32_t)(b1 >> 32) & mkh)));
                accA += fs_term((CT)e.c, (uint32_t)__builtin_popcount(fa));
                accB += fs_term((CT)e.c, (uint32_t)__builtin_popcount(fb));
            }
            const Real S = wave_sum(accA - accB);
            sum = sa == 0 ? (Real)2 * S : (Real)-2 * S;
        } else {
#pragma unroll
            for (int j = 0; j < NS; ++j) {
                if (wok[j]) {
                    const int slot = lane + 64 * j;
                    int na = pA[j], nb = pB[j];
                    if (sA[j] == ia) na = sb;
                    if (sA[j] == ib) na = sa;
                    if (sB[j] == ia) nb = sb;
                    if (sB[j] == ib) nb = sa;
                    char2 va = make_char2(pA[j] ? 1 : -1, na ? 1 : -1);
                    char2 vb = make_char2(pB[j] ? 1 : -1, nb ? 1 : -1);
                    if (sB[j] == ia) vb = make_char2(0, 0);
                    valA[slot] = va;
                    valB[slot] = vb;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            Real acc = (Real)0;
            for (int j = lane; j < Jb0; j += 64) {
                const uint2 e = s_slot[j];
                acc += s_coef[j] * (Real)(entry_delta_bin(valA, e) + entry_delta_bin(valB, e));
            }
            sum = wave_sum(acc);
        }
        if constexpr (sizeof(Real) == 8) return (double)sum * factor;
        else return (double)((float)sum * (float)factor);
    };

    // Group formation (wave 0, one lane per step).  Lanes 0..63 hold the proposals of steps tb..tb+63 and the
    // species at both sites.  A swap of two equal species changes nothing and is rejected, so the group is the run
    // of steps holding the next W different-species swaps (one wave each).  It ends before an equal-species step
    // whose sites an earlier swap of the group may change: every equal-species step inside the group therefore
    // stays one.  The next batch is drawn and its species loaded while the current group is evaluated; the
    // species are then patched with the group's commits, in step order, before the next group is formed.
    int npia = 0, npib = 0, nsa = 0, nsb = 0;
    double npu = 0.0;
    auto draw = [&](int64_t tb) __attribute__((always_inline)) {
        const int64_t t = tb + lane;
        const bool act = t < n_steps;
        if constexpr (INJ) {
            const int64_t tc = act ? t : n_steps - 1;
            npia = (int)inj_ia[inj_base + tc]; npib = (int)inj_ib[inj_base + tc]; npu = inj_u[inj_base + tc];
        } else {
            const Proposal p = draw_proposal_1b(seed, step0 + (uint64_t)t, (uint32_t)r, (uint64_t)n_cells);
            npia = (int)p.ia; npib = (int)p.ib; npu = p.u;
        }
        nsa = act ? (int)occ[(uint32_t)npia] : 0;
        nsb = act ? (int)occ[(uint32_t)npib] : 0;
    };
    // s_pp[k] = the k-th different-species swap of the batch (k < W), the proposal of wave k
    auto form_group = [&](int64_t tb) __attribute__((always_inline)) {
        const bool act = tb + lane < n_steps;
        const bool v = act && nsa != nsb;
        const uint64_t vm = __ballot(v);
        const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(vm >> 32),
                                                        __builtin_amdgcn_mbcnt_lo((uint32_t)vm, 0u));
        if (v && rank < W) s_pp[rank] = SpecProp{npia, npib, npu};
        const int n_act = (int)((n_steps - tb) < (int64_t)SPEC_KP ? (n_steps - tb) : (int64_t)SPEC_KP);
        int end = n_act;
        const uint64_t wth = __ballot(v && rank == W - 1);
        if (wth) end = min(end, __builtin_ctzll(wth) + 1);   // after the W-th different-species swap
        const int nv = min(__popcll(vm), (int)W);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // an equal-species step whose sites meet those of an earlier different-species step ends the group
        // (entries 0 .. rank-1 are the different-species steps before this lane's)
        bool touched = false;
#pragma unroll 4
        for (int k = 0; k < nv; ++k) {
            const SpecProp q = s_pp[k];
            touched |= (k < rank) & ((npia == q.ia) | (npia == q.ib) | (npib == q.ia) | (npib == q.ib));
        }
        const uint64_t tm = __ballot(act && !v && touched && lane < end);
        if (tm) end = min(end, __builtin_ctzll(tm));
        if (lane == 0) {
            s_meta[0] = vm & (end >= 64 ? ~0ull : ((1ull << end) - 1ull));
            s_meta[1] = (unsigned long long)end;
        }
    };
    if (w == 0) {
        draw(0);
        form_group(0);
    }
    __syncthreads();

    for (int64_t t0 = 0; t0 < n_steps;) {
        const uint64_t gvm = s_meta[0];
        const int64_t nstep = (int64_t)s_meta[1];
        const int64_t t1 = t0 + nstep;                 // first step of the next group
        const int nW = __popcll(gvm);                // different-species swaps of the group, one per wave
        const bool active = kw < nW;
        // 1. this wave's swap and its window loads
        int ia = 0, ib = 0;
        double u = 0.0;
        if (active) {
            const SpecProp pp = s_pp[kw];
            ia = pp.ia; ib = pp.ib; u = pp.u;
        }
        ia = __builtin_amdgcn_readfirstlane(ia);
        ib = __builtin_amdgcn_readfirstlane(ib);
        const int a0 = (int)udiv_magic((uint32_t)ia, m12, s12, d12), arem = ia - a0 * (int)d12;
        const int a1 = (int)udiv_magic((uint32_t)arem, m2, sh2, (uint32_t)d2), a2 = arem - a1 * d2;
        const int b0 = (int)udiv_magic((uint32_t)ib, m12, s12, d12), brem = ib - b0 * (int)d12;
        const int b1 = (int)udiv_magic((uint32_t)brem, m2, sh2, (uint32_t)d2), b2 = brem - b1 * d2;
        int sA[NS], sB[NS], pA[NS], pB[NS];
        if (interior(a0, a1, a2)) {
#pragma unroll
            for (int j = 0; j < NS; ++j) sA[j] = wok[j] ? ia + lin[j] : -1;
        } else {
#pragma unroll
            for (int j = 0; j < NS; ++j) {
                int k0 = a0 + wofs[j].y, k1 = a1 + wofs[j].z, k2 = a2 + wofs[j].w;
                k0 -= (k0 >= d0) ? d0 : 0; k1 -= (k1 >= d1) ? d1 : 0; k2 -= (k2 >= d2) ? d2 : 0;
                sA[j] = wok[j] ? (k0 * d1 + k1) * d2 + k2 : -1;
            }
        }
        if (interior(b0, b1, b2)) {
#pragma unroll
            for (int j = 0; j < NS; ++j) sB[j] = wok[j] ? ib + lin[j] : -1;
        } else {
#pragma unroll
            for (int j = 0; j < NS; ++j) {
                int k0 = b0 + wofs[j].y, k1 = b1 + wofs[j].z, k2 = b2 + wofs[j].w;
                k0 -= (k0 >= d0) ? d0 : 0; k1 -= (k1 >= d1) ? d1 : 0; k2 -= (k2 >= d2) ? d2 : 0;
                sB[j] = wok[j] ? (k0 * d1 + k1) * d2 + k2 : -1;
            }
        }
        // every lane loads (a lane without a window slot reads the centre's byte, which shares a line with others),
        // so the loads need no exec-mask branches
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            const int va = (int)occ[(uint32_t)(wok[j] ? sA[j] : ia)], vb = (int)occ[(uint32_t)(wok[j] ? sB[j] : ib)];
            pA[j] = wok[j] ? va : 1;
            pB[j] = wok[j] ? vb : 1;
        }
        if (w == 0 && t1 < n_steps) draw(t1);
        if (lane == 0) s_prop[kw] = make_int2(ia, ib);
        lds_barrier();
        // 2. conflicts with the earlier steps of the group, centres, near flag, energy change.  An earlier step's
        // site can be a slot of this step's windows only inside their offset box: one lane-parallel box test over
        // the 2 kw earlier sites, then an exact comparison of the slots for the steps that pass it (rare).
        int conflict = 0;
        if (kw > 0) {
            bool inb = false;
            if (lane < 2 * kw) {
                const int2 pq = s_prop[lane >> 1];
                int y0, y1, y2;
                coords((lane & 1) ? pq.y : pq.x, y0, y1, y2);
                inb = in_box(y0, y1, y2, a0, a1, a2) || in_box(y0, y1, y2, b0, b1, b2);
            }
            uint64_t cand = __ballot(inb);
            while (cand) {
                const int q = __builtin_ctzll(cand) >> 1;
                cand &= ~(3ull << (2 * q));
                const int2 pq = s_prop[q];
                bool hit = false;
#pragma unroll
                for (int j = 0; j < NS; ++j)
                    hit |= (sA[j] == pq.x) | (sA[j] == pq.y) | (sB[j] == pq.x) | (sB[j] == pq.y);
                if (__ballot(hit) != 0ull) conflict |= 1 << q;
            }
        }
        // near: a slot other than a window's centre is one of the two sites (exact test only when the other site
        // is inside the window's box, or the windows hold images of their centres)
        bool near = false;
        if (self_image || in_box(b0, b1, b2, a0, a1, a2) || in_box(a0, a1, a2, b0, b1, b2)) {
            bool nr = false;
#pragma unroll
            for (int j = 0; j < NS; ++j)
                if (wok[j] && (lane + 64 * j) > 0) nr |= (sA[j] == ia) | (sA[j] == ib) | (sB[j] == ib) | (sB[j] == ia);
            near = __ballot(nr) != 0ull;
        }
        const int sa = __builtin_amdgcn_readfirstlane(pA[0]), sb = __builtin_amdgcn_readfirstlane(pB[0]);
        const bool valid = active && sa != sb;
        double dE = 0.0, lu = 0.0;
        if (valid) dE = evaluate(ia, ib, sA, sB, pA, pB, sa, sb, near);
        // also for a null proposal: a re-evaluation after an earlier commit can make it a real swap
        if (active) {
            if constexpr (sizeof(Real) == 8) lu = log(u);
            else lu = (double)logf((float)u);
        }
        if (lane == 0) {
            SpecRec rec;
            rec.dE = dE; rec.lu = lu; rec.sa = sa; rec.sb = sb; rec.valid = valid; rec.conflict = conflict;
            s_rec[kw] = rec;
        }
        // wave 0 finishes its evaluation first: the next group is formed here from the species as loaded.  A
        // commit of this group is a swap of one of its proposals, so unless a site of the next batch is one of
        // their sites (rare) no commit can change it and the group stands.
        bool spec_hit = false;
        if (w == 0 && t1 < n_steps) {
            form_group(t1);
            bool h = false;
            for (int q = 0; q < nW; ++q) {
                const int2 pq = s_prop[q];
                h |= (npia == pq.x) | (npia == pq.y) | (npib == pq.x) | (npib == pq.y);
            }
            spec_hit = __ballot(h) != 0ull;
        }
        lds_barrier();
        // 3. decisions.  Lane i of every wave holds step i's record; the tentative decisions are taken in parallel,
        // then every step whose windows hold a site changed by an accepted earlier step of the group is
        // re-evaluated on the patched species by its own wave, in step order (all waves follow the same uniform
        // sequence of barriers).
        SpecRec ri = s_rec[lane < nW ? lane : 0];
        auto accepts = [&](const SpecRec &q) __attribute__((always_inline)) -> bool {
            if (!q.valid) return false;
            if constexpr (sizeof(Real) == 8) return q.lu < -beta * q.dE;
            else return (float)q.lu < -(beta_f * (float)q.dE);
        };
        uint64_t accm = __ballot(lane < nW && accepts(ri));
        uint64_t need = __ballot(lane < nW && (ri.conflict & (int)accm) != 0);
        while (need != 0ull) {
            const int f = __builtin_ctzll(need);
            if (tid == 0) {                            // commit list of the accepted steps before f, in order
                int n = 0;
                for (int i = 0; i < f; ++i)
                    if ((accm >> i) & 1ull) {
                        const int2 pi = s_prop[i];
                        const SpecRec q = s_rec[i];
                        s_commit[n++] = make_int4(pi.x, pi.y, q.sb, q.sa);
                    }
            }
            lds_barrier();
            if (kw == f) {
                const int n_commit = __popcll(accm & ((1ull << f) - 1ull));
                int qA[NS], qB[NS];
#pragma unroll
                for (int j = 0; j < NS; ++j) { qA[j] = pA[j]; qB[j] = pB[j]; }
                for (int c = 0; c < n_commit; ++c) {
                    const int4 cm = s_commit[c];
#pragma unroll
                    for (int j = 0; j < NS; ++j) {
                        if (sA[j] == cm.x) qA[j] = cm.z;
                        if (sA[j] == cm.y) qA[j] = cm.w;
                        if (sB[j] == cm.x) qB[j] = cm.z;
                        if (sB[j] == cm.y) qB[j] = cm.w;
                    }
                }
                const int sa2 = __builtin_amdgcn_readfirstlane(qA[0]), sb2 = __builtin_amdgcn_readfirstlane(qB[0]);
                const bool valid2 = sa2 != sb2;
                double dE2 = 0.0;
                if (valid2) dE2 = evaluate(ia, ib, sA, sB, qA, qB, sa2, sb2, near);
                if (lane == 0) {
                    SpecRec r2;
                    r2.dE = dE2; r2.lu = lu; r2.sa = sa2; r2.sb = sb2; r2.valid = valid2; r2.conflict = conflict;
                    s_rec[f] = r2;
                }
            }
            lds_barrier();
            if (lane == f) ri = s_rec[f];
            const uint64_t bit = 1ull << f;
            accm = (accm & ~bit) | (__ballot(lane == f && accepts(ri)) & bit);
            need = __ballot(lane > f && lane < nW && (ri.conflict & (int)accm) != 0);
        }
        // 4. wave 0: occupation writes of the accepted steps (in parallel unless two of them share a site, which
        // only happens through a re-evaluated step: then in step order), then the next group.  The last wave
        // (every wave holds the final records and decisions): running total, statistics, trajectory.
        if (w == 0) {
            const bool mine = (accm >> lane) & 1ull;
            const int2 pl = s_prop[lane < nW ? lane : 0];
            if (__ballot(mine && (ri.conflict & (int)accm) != 0) == 0ull) {
                if (mine) {
                    occ[(uint32_t)pl.x] = (int8_t)ri.sb;
                    occ[(uint32_t)pl.y] = (int8_t)ri.sa;
                }
            } else if (lane == 0) {
                for (int i = 0; i < nW; ++i)
                    if ((accm >> i) & 1ull) {
                        const int2 pi = s_prop[i];
                        const SpecRec q = s_rec[i];
                        occ[pi.x] = (int8_t)q.sb;
                        occ[pi.y] = (int8_t)q.sa;
                    }
            }
            if (t1 < n_steps && spec_hit) {
                // the group's commits in step order, then every lane of the next batch patches its species
                const int crank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(accm >> 32),
                                                                 __builtin_amdgcn_mbcnt_lo((uint32_t)accm, 0u));
                if (mine) s_commit[crank] = make_int4(pl.x, pl.y, ri.sb, ri.sa);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                const int nc = __popcll(accm);
                bool hit = false;
#pragma unroll 4
                for (int c = 0; c < nc; ++c) {
                    const int4 e = s_commit[c];
                    hit |= (npia == e.x) | (npia == e.y) | (npib == e.x) | (npib == e.y);
                    nsa = npia == e.x ? e.z : (npia == e.y ? e.w : nsa);
                    nsb = npib == e.x ? e.z : (npib == e.y ? e.w : nsb);
                }
                if (__ballot(hit) != 0ull) form_group(t1);
            }
        }
        if (w == W - 1) {
            nacc += __popcll(accm);
            // lane i = step i of the group: its energy change (the k-th swap's, if accepted), the running total
            // after it by an inclusive prefix sum, then minimum, moment sums and trajectory samples from the lanes
            const int kl = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(gvm >> 32),
                                                          __builtin_amdgcn_mbcnt_lo((uint32_t)gvm, 0u));
            const bool in = lane < (int)nstep;
            const bool acc_i = in && ((gvm >> lane) & 1ull) && ((accm >> kl) & 1ull);
            const double dk = __shfl(ri.dE, kl < 64 ? kl : 0, 64);
            const double Ei = E + wave_prefix_sum(acc_i ? dk : 0.0);
            emin = fmin(emin, wave_min(in ? Ei : E));
            if (accumulate) {
                const double x = in ? Ei - eref : 0.0;
                s1 += wave_sum(x);
                s2 += wave_sum(x * x);
            }
            while (traj && next_sample <= t0 + nstep) {
                const double Es = readlane_f64(Ei, (int)(next_sample - t0 - 1));
                if (lane == 0) traj[n_sampled * n_rep + slot_r] = Es;
                next_sample += sample_interval;
                ++n_sampled;
            }
            E = readlane_f64(Ei, (int)nstep - 1);
        }
        __syncthreads();                               // occupation writes and the next group
        t0 = t1;
    }
    if (w == W - 1 && lane == 0) {
        job.energy[r] = E;
        job.e_min[r] = emin;
        job.n_accept[r] += nacc;
        job.step[r] = (int64_t)(step0 + (uint64_t)n_steps);
        if (job.accumulate) {
            job.s1[r] += s1;
            job.s2[r] += s2;
        }
    }
}

// ───────────────────────────────────────────── single delta kernel ────────────────────────────────
template <typename Real, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void ce_delta_kernel(const ModelDev *__restrict__ models, int model,
                                                              const int8_t *__restrict__ occ_all, int64_t N,
                                                              const int64_t *__restrict__ ia_all,
                                                              const int64_t *__restrict__ ib_all, double *dE_out) {
    constexpr int BS = 64 * WAVES;
    extern __shared__ __align__(16) unsigned char smem[];
    const ModelDev m = models[model];
    const ChunkLds L = chunk_lds_layout<Real>(m.nB, m.J, m.Wmax, m.Wcap, WAVES);
    uint2 *s_slot = (uint2 *)(smem + L.slot);
    Real *s_coef = (Real *)(smem + L.coef);
    int4 *s_win = (int4 *)(smem + L.win);
    Real2<Real> *valA = (Real2<Real> *)(smem + L.valA);
    Real2<Real> *valB = (Real2<Real> *)(smem + L.valB);
    Real *red = (Real *)(smem + L.red);
    load_model_lds<Real, BS>(m, s_slot, s_coef, s_win, valA, valB);
    __syncthreads();
    const int r = blockIdx.x;
    const NCE_GLOBAL int8_t *occ = gptr(occ_all) + (int64_t)r * N;
    const int64_t ia = ia_all[r], ib = ib_all[r];
    const int sa = occ[ia], sb = occ[ib];
    fill_values<Real, BS>(m, s_win, valA, valB, occ, ia, ib, sa, sb);
    __syncthreads();
    const Real part = entries_partial<Real, BS>(m, s_slot, s_coef, valA, valB, (int)(ia % m.nB), (int)(ib % m.nB));
    const Real sum = block_sum<Real, WAVES>(part, red);
    if (threadIdx.x == 0) {
        if constexpr (sizeof(Real) == 8) dE_out[r] = (double)sum * m.factor;
        else dE_out[r] = (double)((float)sum * (float)m.factor);
    }
}

// ───────────────────────────────────────────── energy kernels ─────────────────────────────────────
constexpr int ENERGY_BS = 256;

__global__ __launch_bounds__(ENERGY_BS) void ce_energy_kernel(const ModelDev *__restrict__ models, int model,
                                                              const int8_t *__restrict__ occ_all, int64_t N,
                                                              double *__restrict__ partial) {
    extern __shared__ __align__(16) unsigned char smem[];
    const ModelDev m = models[model];
    const int nrm = m.Rt * m.mo;
    int4 *s_rows = (int4 *)smem;                        // x = basis | (sublattice * 2 + bf) << 16
    double *s_coef = (double *)(smem + align16((size_t)nrm * sizeof(int4)));
    double *s_pf = s_coef + m.Rt;                      // [n_sl][2][2]
    double *s_red = s_pf + m.n_sl * 4;
    for (int i = threadIdx.x; i < nrm; i += ENERGY_BS) {
        int4 e = m.rows[i];
        e.x = e.x | ((m.basis_to_sl[e.x] * 2 + m.row_bf[i]) << 16);
        s_rows[i] = e;
    }
    for (int i = threadIdx.x; i < m.Rt; i += ENERGY_BS) s_coef[i] = m.row_coef[i];
    for (int i = threadIdx.x; i < m.n_sl * 4; i += ENERGY_BS) s_pf[i] = m.pf[i];
    __syncthreads();
    const int r = blockIdx.y;
    const int8_t *occ = occ_all + (int64_t)r * N;
    const int64_t cell = (int64_t)blockIdx.x * ENERGY_BS + threadIdx.x;
    double acc = 0.0;
    if (cell < m.n_cells) {
        const Coord c = site_coord(m, cell * m.nB);
        for (int row = 0; row < m.Rt; ++row) {
            double p = 1.0;
            for (int q = 0; q < m.mo; ++q) {
                int4 e = s_rows[row * m.mo + q];
                const int pfo = (e.x >> 16) * 2;
                e.x &= 0xffff;
                const int64_t s = window_site(m, c, e);
                p *= s_pf[pfo + occ[s]];
            }
            acc += s_coef[row] * p;
        }
    }
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < ENERGY_BS / 64; ++w) s += s_red[w];
        partial[(int64_t)r * gridDim.x + blockIdx.x] = s;
    }
}

// Full energies of the binary fast path, one wave per site at a time.  Every cluster is counted once, by the entry
// centred on its first member: E = zerolet + factor * sum_x (-1)^[s_x = 0] * sum_e |c_e| (-1)^parity(bits_x & m_e)
// over those entries, the window-parity sum of the chunk kernel (window bits [species == 0], the centre bit set as
// the carrier of each coefficient's sign).  Each block takes EF_SITES consecutive sites, 64 per wave: consecutive
// windows overlap in the caches, and the cell coordinate is stepped rather than divided.  The next site's window
// is loaded while the current one is summed; the centre's species is bit 0 of its window (slot 0 is the site), and
// each lane keeps its own signed sum over the wave's sites, reduced once at the end.  Same partial layout as
// ce_energy_kernel (blocks of 256 cells, nB = 1).
constexpr int EF_BS = 256;
constexpr int EF_SITES = 256;
static_assert(EF_SITES == ENERGY_BS, "partial sums are shared with ce_energy_finish_kernel");

__global__ __launch_bounds__(EF_BS) void ce_energy_fast_kernel(const ModelDev *__restrict__ models, int model,
                                                               const int8_t *__restrict__ occ_all, int64_t N,
                                                               double *__restrict__ partial) {
    extern __shared__ __align__(16) unsigned char smem[];
    const ModelDev &mg = models[model];
    const int nlo = mg.fse_nlo, nfull = mg.fse_nfull, ntail = mg.fse_ntail, Wn = mg.W_b[0];
    const int d0 = mg.d0, d1 = mg.d1, d2 = mg.d2;
    const size_t fsb = fs_bytes<double>(nlo, nfull, ntail);
    {
        const uint4 *src = (const uint4 *)mg.fs_tab_e;
        uint4 *dst = (uint4 *)smem;
        for (int i = threadIdx.x; i < (int)(fsb / 16); i += EF_BS) dst[i] = src[i];
    }
    double *s_red = (double *)(smem + align16(fsb));
    const FsLo<double> *s_lo = (const FsLo<double> *)smem;
    const FsFull *s_full = (const FsFull *)(s_lo + 64 * nlo);
    const FsTail *s_tail = (const FsTail *)(s_full + 64 * nfull);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int4 wofs[3];
    bool wok[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const int slot = lane + 64 * j;
        wok[j] = slot < Wn;
        wofs[j] = wok[j] ? mg.win[slot] : make_int4(0, 0, 0, 0);
    }
    __syncthreads();
    const NCE_GLOBAL int8_t *occ = gptr(occ_all) + (int64_t)blockIdx.y * N;
    const int64_t x0 = (int64_t)blockIdx.x * EF_SITES + 64 * w;
    const int n_here = x0 < N ? (int)(N - x0 < 64 ? N - x0 : 64) : 0;
    Coord c = site_coord(mg, x0 < N ? x0 : 0);
    int sp_next[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) sp_next[j] = (wok[j] && n_here > 0) ? occ[window_site(mg, c, wofs[j])] : 1;
    double acc = 0.0;
    for (int i = 0; i < n_here; ++i) {
        unsigned long long Wd[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) Wd[j] = __ballot(wok[j] && sp_next[j] == 0);
        if (++c.k2 == d2) {                            // next site (one basis): step the cell coordinate
            c.k2 = 0;
            if (++c.k1 == d1) {
                c.k1 = 0;
                if (++c.k0 == d0) c.k0 = 0;
            }
        }
        if (i + 1 < n_here) {
#pragma unroll
            for (int j = 0; j < 3; ++j)
                if (wok[j]) sp_next[j] = occ[window_site(mg, c, wofs[j])];
        }
        const bool centre0 = (Wd[0] & 1ull) != 0;      // slot 0 is the site itself
        Wd[0] |= 1ull;
        double a = 0.0;
        const uint32_t w32 = (uint32_t)Wd[0], whi = (uint32_t)(Wd[0] >> 32);
        for (int it = 0; it < nlo; ++it) {
            const FsLo<double> e = s_lo[64 * it + lane];
            a += fs_term((uint64_t)e.c, (uint32_t)__builtin_popcount(w32 & e.m));
        }
        for (int it = 0; it < nfull; ++it) {
            const FsFull e = s_full[64 * it + lane];
            a += fs_term((uint64_t)e.c,
                         (uint32_t)__builtin_popcount(and_xor(w32, (uint32_t)e.m, whi & (uint32_t)(e.m >> 32))));
        }
        for (int it = 0; it < ntail; ++it) {
            const FsTail e = s_tail[64 * it + lane];
            const unsigned long long w1 = e.word == 1 ? Wd[1] : Wd[2];
            const uint32_t f = and_xor(w32, (uint32_t)e.m0, and_xor(whi, (uint32_t)(e.m0 >> 32),
                                       and_xor((uint32_t)w1, (uint32_t)e.mk, (uint32_t)(w1 >> 32) & (uint32_t)(e.mk >> 32))));
            a += fs_term((uint64_t)e.c, (uint32_t)__builtin_popcount(f));
        }
        acc += centre0 ? -a : a;
    }
    acc = wave_sum(acc);
    if (lane == 0) s_red[w] = acc;
    __syncthreads();
    if (threadIdx.x == 0)
        partial[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = (s_red[0] + s_red[1]) + (s_red[2] + s_red[3]);
}

__global__ void ce_energy_finish_kernel(const ModelDev *__restrict__ models, int model,
                                        const double *__restrict__ partial, int n_part, double *E) {
    const ModelDev m = models[model];
    const int r = blockIdx.x;
    double acc = 0.0;
    for (int i = threadIdx.x; i < n_part; i += 64) acc += partial[(int64_t)r * n_part + i];
    acc = wave_sum(acc);
    if (threadIdx.x == 0) E[r] = m.zerolet + m.factor * acc;
}

// ───────────────────────────────────────────── exchange kernel ────────────────────────────────────
// One workgroup per job.  The pairs of one parity are disjoint, so each lane takes pairs p = lane, lane + 256, ...
// with its own Philox uniform (counter (ex_index, p)); the new slot maps are then rewritten in parallel.
constexpr int EXCH_BS = 256;

__global__ __launch_bounds__(EXCH_BS) void pt_exchange_kernel(const nce_pt_exchange_job *__restrict__ jobs) {
    __shared__ int s_ns;
    const nce_pt_exchange_job j = jobs[blockIdx.x];
    const int R = j.n_replicas;
    const uint32_t k0 = (uint32_t)j.seed, k1 = (uint32_t)(j.seed >> 32);
    // exchange counter from the device when the job gives one (captured cycles), else from the job table
    const uint64_t ex_index = j.ex_count ? (uint64_t)(j.ex_count[0] + 1) : j.ex_index;
    int parity;
    if (j.inj_parity >= 0) {
        parity = j.inj_parity & 1;
    } else {
        uint32_t z[4] = {(uint32_t)ex_index, (uint32_t)(ex_index >> 32), 0xffffffffu, 2u};
        philox4x32_10(z, k0, k1);
        parity = (int)(z[0] & 1u);
    }
    if (threadIdx.x == 0) s_ns = 0;
    __syncthreads();
    int ns = 0;
    for (int p = threadIdx.x; 2 * p <= R - 2; p += EXCH_BS) {
        const int a = 2 * p + parity, b = a + 1;
        if (b >= R) continue;
        if (j.ladder_id && j.ladder_id[a] != j.ladder_id[b]) continue;
        double u;
        if (j.inj_u) {
            u = j.inj_u[p];
        } else {
            uint32_t w[4] = {(uint32_t)ex_index, (uint32_t)(ex_index >> 32), (uint32_t)p, 3u};
            philox4x32_10(w, k0, k1);
            u = ((double)((((uint64_t)w[1] << 32) | w[0]) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        }
        const int ra = j.rep_of_slot[a], rb = j.rep_of_slot[b];
        const double la = (j.beta_slot[b] - j.beta_slot[a]) * (j.energy[rb] - j.energy[ra]);
        if (log(u) < la) {
            j.rep_of_slot[a] = rb;
            j.rep_of_slot[b] = ra;
            ++ns;
        }
    }
    if (ns) atomicAdd(&s_ns, ns);
    __syncthreads();
    for (int t = threadIdx.x; t < R; t += EXCH_BS) {
        const int rep = j.rep_of_slot[t];
        j.beta_rep[rep] = j.beta_slot[t];
        j.slot_of_rep[rep] = t;
        if (j.traj_last) j.traj_last[t] = j.energy[rep];
    }
    if (threadIdx.x == 0) {
        j.n_swaps[0] += s_ns;
        if (j.ex_count) j.ex_count[0] = (int64_t)ex_index;     // every thread read it before the barrier above
    }
}

// Moments of a chunk folded into the per-slot accumulators (Chan, Golub, LeVeque: pairwise update of count, mean and
// sum of squared deviations), one thread per (job, replica).  slot_of_rep is a permutation within a job, so every
// slot is written by exactly one thread.  Runs after the chunk and before the exchange that may move the slots.
constexpr int MERGE_BS = 256;

__global__ __launch_bounds__(MERGE_BS) void pt_merge_kernel(const nce_pt_job *__restrict__ jobs,
                                                            const int2 *__restrict__ items, int n_items,
                                                            int64_t n_steps) {
    const int i = blockIdx.x * MERGE_BS + threadIdx.x;
    if (i >= n_items) return;
    const int2 it = items[i];
    const nce_pt_job j = jobs[it.x];
    if (!j.acc_n || !j.accumulate) return;
    const int r = it.y;
    const int t = j.slot_of_rep[r];
    const double n = (double)n_steps;
    const double s1 = j.s1[r], s2 = j.s2[r];
    const double mean_c = __dadd_rn(j.e_ref[r], s1 / n);
    const double m2_c = fmax(__dadd_rn(s2, -__dmul_rn(s1, s1) / n), 0.0);
    const double na = j.acc_n[t], tot = __dadd_rn(na, n);
    const double delta = __dadd_rn(mean_c, -j.acc_mean[t]);
    j.acc_mean[t] = __dadd_rn(j.acc_mean[t], __dmul_rn(delta, n) / tot);
    j.acc_m2[t] = __dadd_rn(j.acc_m2[t], __dadd_rn(m2_c, __dmul_rn(__dmul_rn(__dmul_rn(delta, delta), na), n) / tot));
    j.acc_n[t] = tot;
    j.s1[r] = 0.0;
    j.s2[r] = 0.0;
    const_cast<double *>(j.e_ref)[r] = j.energy[r];
}

// ───────────────────────────────────────────── host side ──────────────────────────────────────────
extern "C" int nce_ctx_create(int32_t device_ordinal, nce_ctx **out) {
    if (!out) return set_err(NCE_ERR_ARG, "out is NULL");
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device_ordinal < 0 || device_ordinal >= n)
        return set_err(NCE_ERR_ARG, "device ordinal " + std::to_string(device_ordinal) + " out of range (" +
                                        std::to_string(n) + " devices)");
    HIP_TRY(hipSetDevice(device_ordinal));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device_ordinal));
    double *noalloc = nullptr;                    // with this device current: its copy of the no-allocation scratch
    HIP_TRY(hipGetSymbolAddress((void **)&noalloc, HIP_SYMBOL(scratch)));
    nce_ctx *c = new nce_ctx();
    c->device = device_ordinal;
    c->noalloc_partial = noalloc;
    c->n_cu = prop.multiProcessorCount;
    c->max_lds = prop.maxSharedMemoryPerMultiProcessor ? prop.maxSharedMemoryPerMultiProcessor : 65536;
    *out = c;
    return NCE_OK;
}

extern "C" void nce_ctx_destroy(nce_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipDeviceSynchronize();
    for (void *p : c->allocs) (void)hipFree(p);
    if (c->d_models) (void)hipFree(c->d_models);
    for (nce_ctx::Partial &pb : c->partials)
        if (pb.p) (void)hipFree(pb.p);
    for (nce_ctx::Partial &pb : c->h0_rows)
        if (pb.p) (void)hipFree(pb.p);
    for (nce_ctx::Partial &pb : c->split_mem)
        if (pb.p) (void)hipFree(pb.p);
    if (c->d_nn_models) (void)hipFree(c->d_nn_models);
    for (nce_ctx::StageRing *ring : c->rings) {
        for (auto &st : ring->stage) {
            if (st.host) (void)hipHostFree(st.host);
            if (st.dev) (void)hipFree(st.dev);
            if (st.ev) (void)hipEventDestroy(st.ev);
        }
        delete ring;
    }
    delete c;
}

template <typename T>
static int upload(nce_ctx *c, const T *host, size_t n, const T **dev) {
    if (n == 0) { *dev = nullptr; return NCE_OK; }
    if (!host) return set_err(NCE_ERR_ARG, "missing host array");
    void *p = nullptr;
    HIP_TRY(hipMalloc(&p, n * sizeof(T)));
    c->allocs.push_back(p);
    HIP_TRY(hipMemcpy(p, host, n * sizeof(T), hipMemcpyHostToDevice));
    *dev = (const T *)p;
    return NCE_OK;
}

#define UP(h, n, d)                                                                                    \
    do {                                                                                               \
        int rc_ = upload(c, h, n, d);                                                                  \
        if (rc_) return rc_;                                                                           \
    } while (0)

static int sync_models(nce_ctx *c) {
    if ((int)c->models.size() > c->cap_models) {
        int cap = std::max(16, 2 * (int)c->models.size());
        void *p = nullptr;
        HIP_TRY(hipMalloc(&p, cap * sizeof(ModelDev)));
        if (c->d_models) HIP_TRY(hipFree(c->d_models));
        c->d_models = (ModelDev *)p;
        c->cap_models = cap;
    }
    HIP_TRY(hipMemcpy(c->d_models, c->models.data(), c->models.size() * sizeof(ModelDev), hipMemcpyHostToDevice));
    return NCE_OK;
}

// magic numbers for unsigned division by d (1 <= d < 2^31) valid for every 32-bit dividend
static void magic_u32(uint32_t d, uint32_t *m, uint32_t *sh) {
    uint32_t l = 0;
    while (((uint64_t)1 << l) < d) ++l;                              // l = ceil(log2 d)
    const uint64_t mm = ((((uint64_t)1 << l) - d) << 32) / d + 1;
    *m = (uint32_t)mm;
    *sh = l == 0 ? 0 : l - 1;
    if (l == 0) { *m = 0; *sh = 0; }
}

// Fast-path tables: entries scheduled in iterations of 64*waves, one second window word per iteration.
static int build_fast_tables(nce_ctx *c, const nce_ce_model_desc *d, ModelDev &m) {
    m.fs_ok = 0;
    m.fs_nlo = m.fs_nfull = m.fs_ntail = 0;
    m.fse_nlo = m.fse_nfull = m.fse_ntail = 0;
    m.fs_tab32 = m.fs_tab64 = m.fs_tab_e = nullptr;
    if (d->n_basis != 1 || d->n_active != 1 || d->active_nbasis[0] != 1 || d->max_window > 192) return NCE_OK;
    for (int sl = 0; sl < d->n_sublattices; ++sl) {
        const double *pf = d->pf + sl * 4;
        if (pf[0] != 1.0 || pf[1] != 1.0 || pf[2] != -1.0 || pf[3] != 1.0) return NCE_OK;
    }
    // classify the entries of the one basis site (entries with a zero coefficient contribute nothing)
    struct Ent {
        unsigned long long m0, mk;
        double c;
        int word;
    };
    // lo / full / tail: every entry (swap energy changes); the *_e lists: the entries whose centre is member 0 of
    // their cluster (members are listed in the cluster's own order, so that member sits in window slot 0), one
    // per cluster and translation, which is what a full energy sums
    std::vector<Ent> lo, full, tail, lo_e, full_e, tail_e;
    const int Jb = d->entries_per_basis[0];
    for (int j = 0; j < Jb; ++j) {
        Ent e{0ull, 0ull, d->entry_coef[j], 0};
        int centre = 0;
        for (int q = 0; q < 4; ++q) {
            const uint32_t sl = d->entry_slot[(size_t)j * 4 + q];
            if (sl == 0xFFFFu) continue;
            if (sl == 0) { ++centre; continue; }
            const int word = (int)(sl >> 6);
            const unsigned long long bit = 1ull << (sl & 63);
            if (word == 0) e.m0 ^= bit;
            else {
                if (e.word && e.word != word) return NCE_OK;   // members in two upper words: general path only
                e.word = word;
                e.mk ^= bit;
            }
        }
        if (centre != 1) return NCE_OK;
        if (e.c == 0.0) continue;
        if (e.c < 0.0) { e.m0 |= 1ull; e.c = -e.c; }
        const bool first = d->entry_slot[(size_t)j * 4] == 0;
        if (e.word) {
            tail.push_back(e);
            if (first) tail_e.push_back(e);
        } else if ((e.m0 >> 32) == 0) {
            lo.push_back(e);
            if (first) lo_e.push_back(e);
        } else {
            full.push_back(e);
            if (first) full_e.push_back(e);
        }
    }
    const int nlo = (int)((lo.size() + 63) / 64), nfull = (int)((full.size() + 63) / 64),
              ntail = (int)((tail.size() + 63) / 64);
    const int nlo_e = (int)((lo_e.size() + 63) / 64), nfull_e = (int)((full_e.size() + 63) / 64),
              ntail_e = (int)((tail_e.size() + 63) / 64);
    std::vector<unsigned char> t32(fs_bytes<float>(nlo, nfull, ntail), 0), t64(fs_bytes<double>(nlo, nfull, ntail), 0),
        te(fs_bytes<double>(nlo_e, nfull_e, ntail_e), 0);
    auto bits32 = [](double v) { float f = (float)v; uint32_t u; std::memcpy(&u, &f, 4); return u; };
    auto bits64 = [](double v) { uint64_t u; std::memcpy(&u, &v, 8); return u; };
    {
        FsLo<float> *l32 = (FsLo<float> *)t32.data();
        FsLo<double> *l64 = (FsLo<double> *)t64.data();
        for (size_t i = 0; i < lo.size(); ++i) {
            l32[i].m = l64[i].m = (uint32_t)lo[i].m0;
            l32[i].c = bits32(lo[i].c);
            l64[i].c = bits64(lo[i].c);
        }
        FsFull *f32 = (FsFull *)(l32 + 64 * nlo), *f64 = (FsFull *)(l64 + 64 * nlo);
        for (size_t i = 0; i < full.size(); ++i) {
            f32[i].m = f64[i].m = full[i].m0;
            f32[i].c = bits32(full[i].c);
            f64[i].c = bits64(full[i].c);
        }
        FsTail *u32 = (FsTail *)(f32 + 64 * nfull), *u64 = (FsTail *)(f64 + 64 * nfull);
        for (size_t i = 0; i < tail.size(); ++i) {
            u32[i].m0 = u64[i].m0 = tail[i].m0;
            u32[i].mk = u64[i].mk = tail[i].mk;
            u32[i].word = u64[i].word = (uint32_t)tail[i].word;
            u32[i].c = bits32(tail[i].c);
            u64[i].c = bits64(tail[i].c);
        }
        for (size_t i = tail.size(); i < (size_t)64 * ntail; ++i) u32[i].word = u64[i].word = 1;
        // full-energy table, FP64 layout over the first-member entries
        FsLo<double> *le = (FsLo<double> *)te.data();
        for (size_t i = 0; i < lo_e.size(); ++i) {
            le[i].m = (uint32_t)lo_e[i].m0;
            le[i].c = bits64(lo_e[i].c);
        }
        FsFull *fe = (FsFull *)(le + 64 * nlo_e);
        for (size_t i = 0; i < full_e.size(); ++i) {
            fe[i].m = full_e[i].m0;
            fe[i].c = bits64(full_e[i].c);
        }
        FsTail *ue = (FsTail *)(fe + 64 * nfull_e);
        for (size_t i = 0; i < tail_e.size(); ++i) {
            ue[i].m0 = tail_e[i].m0;
            ue[i].mk = tail_e[i].mk;
            ue[i].word = (uint32_t)tail_e[i].word;
            ue[i].c = bits64(tail_e[i].c);
        }
        for (size_t i = tail_e.size(); i < (size_t)64 * ntail_e; ++i) ue[i].word = 1;
    }
    if (!t32.empty()) {
        UP(t32.data(), t32.size(), &m.fs_tab32);
        UP(t64.data(), t64.size(), &m.fs_tab64);
    }
    if (!te.empty()) UP(te.data(), te.size(), &m.fs_tab_e);
    m.fs_nlo = nlo; m.fs_nfull = nfull; m.fs_ntail = ntail;
    m.fse_nlo = nlo_e; m.fse_nfull = nfull_e; m.fse_ntail = ntail_e;
    m.fs_ok = 1;
    return NCE_OK;
}

extern "C" int nce_ce_model_create(nce_ctx *c, const nce_ce_model_desc *d, int32_t *model_id) {
    if (!c || !d || !model_id) return set_err(NCE_ERR_ARG, "NULL argument");
    HIP_TRY(hipSetDevice(c->device));
    if (d->n_basis < 1 || d->max_order < 1 || d->max_order > 4 || d->n_entries < 1 || d->max_window < 1 ||
        d->n_rows < 1 || d->n_active < 1 || d->max_active_basis < 1)
        return set_err(NCE_ERR_ARG, "model sizes out of range (max_order must be 1..4)");
    for (int i = 0; i < 3; ++i)
        if (d->d[i] < 1) return set_err(NCE_ERR_ARG, "Smith normal form diagonal must be positive");
    if ((int64_t)d->d[0] * d->d[1] * d->d[2] != d->n_cells) return set_err(NCE_ERR_ARG, "n_cells != d0*d1*d2");
    ModelDev m{};
    m.nB = d->n_basis; m.J = d->n_entries; m.Wmax = d->max_window; m.mo = d->max_order; m.Rt = d->n_rows;
    m.n_act = d->n_active; m.max_act_b = d->max_active_basis; m.n_sl = d->n_sublattices;
    m.d0 = d->d[0]; m.d1 = d->d[1]; m.d2 = d->d[2];
    m.n_cells = d->n_cells; m.n_sites = d->n_cells * d->n_basis;
    m.factor = d->factor; m.zerolet = d->zerolet;
    int Wcap = 1;
    while (Wcap < m.Wmax + 1) Wcap <<= 1;
    m.Wcap = Wcap;
    const size_t nJ = (size_t)m.nB * m.J;
    // validate and remap slots (0xFFFF -> Wcap - 1)
    std::vector<uint2> slots(nJ);
    for (size_t i = 0; i < nJ; ++i) {
        uint32_t q[4];
        const int b = (int)(i / m.J);
        for (int k = 0; k < 4; ++k) {
            uint32_t s = d->entry_slot[i * 4 + k];
            if (s == 0xFFFFu) s = (uint32_t)(Wcap - 1);
            else if ((int)s >= d->window_per_basis[b]) return set_err(NCE_ERR_ARG, "entry slot outside its window");
            q[k] = s;
        }
        slots[i] = make_uint2(q[0] | (q[1] << 16), q[2] | (q[3] << 16));
    }
    for (int b = 0; b < m.nB; ++b) {
        if (d->entries_per_basis[b] < 0 || d->entries_per_basis[b] > m.J || d->window_per_basis[b] < 1 ||
            d->window_per_basis[b] > m.Wmax)
            return set_err(NCE_ERR_ARG, "per-basis entry / window counts out of range");
    }
    const size_t nW = (size_t)m.nB * m.Wmax, nR = (size_t)m.Rt * m.mo;
    for (size_t i = 0; i < nW; ++i) {
        const int32_t *w = d->window + 4 * i;
        if (w[0] < 0 || w[0] >= m.nB || w[1] < 0 || w[1] >= m.d0 || w[2] < 0 || w[2] >= m.d1 || w[3] < 0 ||
            w[3] >= m.d2)
            return set_err(NCE_ERR_ARG, "window offset not reduced into the cell");
    }
    for (size_t i = 0; i < nR; ++i) {
        const int32_t *w = d->rows + 4 * i;
        if (w[0] < 0 || w[0] >= m.nB || w[1] < 0 || w[1] >= m.d0 || w[2] < 0 || w[2] >= m.d1 || w[3] < 0 ||
            w[3] >= m.d2 || d->row_bf[i] < 0 || d->row_bf[i] > 1)
            return set_err(NCE_ERR_ARG, "forward row offset not reduced into the cell");
    }
    for (int b = 0; b < m.nB; ++b)
        if (d->basis_to_sl[b] < 0 || d->basis_to_sl[b] >= m.n_sl) return set_err(NCE_ERR_ARG, "basis_to_sl");
    for (int a = 0; a < m.n_act; ++a) {
        if (d->active_nbasis[a] < 1 || d->active_nbasis[a] > m.max_act_b) return set_err(NCE_ERR_ARG, "active_nbasis");
        for (int k = 0; k < d->active_nbasis[a]; ++k) {
            const int b = d->active_basis[a * m.max_act_b + k];
            if (b < 0 || b >= m.nB) return set_err(NCE_ERR_ARG, "active_basis");
        }
    }
    std::vector<float> coef_f(nJ);
    for (size_t i = 0; i < nJ; ++i) coef_f[i] = (float)d->entry_coef[i];
    UP(d->entries_per_basis, (size_t)m.nB, &m.J_b);
    UP(d->window_per_basis, (size_t)m.nB, &m.W_b);
    UP(slots.data(), nJ, &m.ent_slot);
    UP(d->entry_coef, nJ, &m.ent_coef);
    UP(coef_f.data(), nJ, &m.ent_coef_f);
    UP((const int4 *)d->window, nW, &m.win);
    UP(d->pf, (size_t)m.n_sl * 4, &m.pf);
    UP(d->basis_to_sl, (size_t)m.nB, &m.basis_to_sl);
    UP(d->active_basis, (size_t)m.n_act * m.max_act_b, &m.act_b);
    UP(d->active_nbasis, (size_t)m.n_act, &m.act_nb);
    UP(d->active_cum, (size_t)m.n_act, &m.act_cum);
    UP((const int4 *)d->rows, nR, &m.rows);
    UP(d->row_bf, nR, &m.row_bf);
    UP(d->row_coef, (size_t)m.Rt, &m.row_coef);
    magic_u32((uint32_t)(m.d1 * m.d2), &m.div12_m, &m.div12_s);
    magic_u32((uint32_t)m.d2, &m.div2_m, &m.div2_s);
    if ((int64_t)m.d1 * m.d2 >= ((int64_t)1 << 31)) return set_err(NCE_ERR_UNSUPPORTED, "cell too large (d1*d2)");
    if (int rcf = build_fast_tables(c, d, m)) return rcf;
    c->models.push_back(m);
    c->lds_chunk64.push_back(chunk_lds_layout<double>(m.nB, m.J, m.Wmax, m.Wcap, 8).total);
    c->lds_chunk32.push_back(chunk_lds_layout<float>(m.nB, m.J, m.Wmax, m.Wcap, 8).total);
    c->lds_energy.push_back(align16(nR * sizeof(int4)) + (m.Rt + m.n_sl * 4 + ENERGY_BS / 64) * sizeof(double));
    int rc = sync_models(c);
    if (rc) return rc;
    *model_id = (int32_t)c->models.size() - 1;
    return NCE_OK;
}

static int check_model(nce_ctx *c, int32_t model) {
    if (!c) return set_err(NCE_ERR_ARG, "NULL context");
    if (model < 0 || model >= (int32_t)c->models.size()) return set_err(NCE_ERR_ARG, "unknown model id");
    return NCE_OK;
}

// the partial-sum buffer of stream s, grown to `need` doubles (the old buffer is freed after the stream drains)
static int stream_buffer(std::vector<nce_ctx::Partial> &bufs, hipStream_t s, size_t need, double **out) {
    nce_ctx::Partial *pb = nullptr;
    for (nce_ctx::Partial &q : bufs)
        if (q.stream == s) pb = &q;
    if (!pb) {
        bufs.push_back(nce_ctx::Partial());
        pb = &bufs.back();
        pb->stream = s;
    }
    if (need > pb->cap) {
        if (pb->p) {
            HIP_TRY(hipStreamSynchronize(s));
            HIP_TRY(hipFree(pb->p));
            pb->p = nullptr;
            pb->cap = 0;
        }
        HIP_TRY(hipMalloc(&pb->p, need * sizeof(double)));
        pb->cap = need;
    }
    *out = pb->p;
    return NCE_OK;
}

// the stream's buffer of per-block partial sums (doubles), grown as needed
static int partial_for(nce_ctx *c, hipStream_t s, size_t need, double **out) {
    return stream_buffer(c->partials, s, need, out);
}

template <typename K>
static int allow_lds(K kernel, size_t lds) {
    if (lds > 65536)
        HIP_TRY(hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    return NCE_OK;
}

// full energies: argument checks and the number of per-block partial sums per replica
static int ce_energy_blocks(nce_ctx *c, int32_t model, const int8_t *occ, int32_t R, int64_t N, const double *E,
                            int *nblk) {
    int rc = check_model(c, model);
    if (rc) return rc;
    const ModelDev &m = c->models[model];
    if (N != m.n_sites) return set_err(NCE_ERR_ARG, "n_sites does not match the model's cell");
    if (R < 1 || !occ || !E) return set_err(NCE_ERR_ARG, "bad replica count or NULL buffer");
    const int64_t nb = (m.n_cells + ENERGY_BS - 1) / ENERGY_BS;
    if (nb > (1 << 30)) return set_err(NCE_ERR_ARG, "cell too large");
    *nblk = (int)nb;
    return NCE_OK;
}

static int ce_energy_launch(nce_ctx *c, int32_t model, const int8_t *occ, int32_t R, int64_t N, double *E, int nblk,
                            double *partial, hipStream_t s) {
    const ModelDev &m = c->models[model];
    const char *fe = std::getenv("NCE_CE_ENERGY_ROWS");              // tests: the forward-row kernel on any model
    if (m.fs_ok && m.nB == 1 && m.fs_tab_e && !(fe && fe[0] == '1')) {
        const size_t lds = align16(fs_bytes<double>(m.fse_nlo, m.fse_nfull, m.fse_ntail)) + 4 * sizeof(double);
        if (lds > 160 * 1024) return set_err(NCE_ERR_UNSUPPORTED, "entry tables do not fit in LDS");
        if (int rc = allow_lds(ce_energy_fast_kernel, lds)) return rc;
        hipLaunchKernelGGL(ce_energy_fast_kernel, dim3(nblk, R), dim3(EF_BS), lds, s, c->d_models, model, occ, N,
                           partial);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(ce_energy_finish_kernel, dim3(R), dim3(64), 0, s, c->d_models, model, partial, nblk, E);
        HIP_TRY(hipGetLastError());
        return NCE_OK;
    }
    const size_t lds = c->lds_energy[model];
    if (lds > 160 * 1024) return set_err(NCE_ERR_UNSUPPORTED, "forward rows do not fit in LDS");
    if (int rc = allow_lds(ce_energy_kernel, lds)) return rc;
    hipLaunchKernelGGL(ce_energy_kernel, dim3(nblk, R), dim3(ENERGY_BS), lds, s, c->d_models, model, occ, N, partial);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(ce_energy_finish_kernel, dim3(R), dim3(64), 0, s, c->d_models, model, partial, nblk, E);
    HIP_TRY(hipGetLastError());
    return NCE_OK;
}

extern "C" int nce_ce_energy(nce_ctx *c, int32_t model, const int8_t *occ, int32_t R, int64_t N, double *E,
                             void *stream) {
    int nblk = 0;
    if (int rc = ce_energy_blocks(c, model, occ, R, N, E, &nblk)) return rc;
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t s = (hipStream_t)stream;
    double *partial = nullptr;
    if (int rc = partial_for(c, s, (size_t)R * nblk, &partial)) return rc;
    return ce_energy_launch(c, model, occ, R, N, E, nblk, partial, s);
}

extern "C" int nce_ce_energy_noalloc(nce_ctx *c, int32_t model, const int8_t *occ, int32_t R, int64_t N, double *E,
                                     void *stream) {
    int nblk = 0;
    if (int rc = ce_energy_blocks(c, model, occ, R, N, E, &nblk)) return rc;
    if ((size_t)R * (size_t)nblk > (size_t)NCE_NOALLOC_PARTIALS)
        return set_err(NCE_ERR_ARG, "n_replicas * ceil(n_cells / 256) = " + std::to_string((size_t)R * nblk) +
                                        " partial sums exceed the " + std::to_string((size_t)NCE_NOALLOC_PARTIALS) +
                                        " of the no-allocation scratch");
    if (!c->noalloc_partial) return set_err(NCE_ERR_HIP, "no-allocation scratch address unknown");
    HIP_TRY(hipSetDevice(c->device));
    return ce_energy_launch(c, model, occ, R, N, E, nblk, c->noalloc_partial, (hipStream_t)stream);
}

template <typename Real, int W>
static int launch_delta(nce_ctx *c, int32_t model, const int8_t *occ, int32_t R, int64_t N, const int64_t *ia,
                        const int64_t *ib, double *dE, hipStream_t s) {
    const ModelDev &m = c->models[model];
    const size_t lds = chunk_lds_layout<Real>(m.nB, m.J, m.Wmax, m.Wcap, W).total;
    if (lds > 160 * 1024) return set_err(NCE_ERR_UNSUPPORTED, "template does not fit in LDS");
    if (int rc = allow_lds(ce_delta_kernel<Real, W>, lds)) return rc;
    hipLaunchKernelGGL((ce_delta_kernel<Real, W>), dim3(R), dim3(64 * W), lds, s, c->d_models, model, occ, N, ia, ib,
                       dE);
    HIP_TRY(hipGetLastError());
    return NCE_OK;
}

extern "C" int nce_ce_delta(nce_ctx *c, int32_t model, const int8_t *occ, int32_t R, int64_t N, const int64_t *ia,
                            const int64_t *ib, double *dE, int32_t precision, void *stream) {
    int rc = check_model(c, model);
    if (rc) return rc;
    if (N != c->models[model].n_sites) return set_err(NCE_ERR_ARG, "n_sites does not match the model's cell");
    if (R < 1 || !occ || !ia || !ib || !dE) return set_err(NCE_ERR_ARG, "bad replica count or NULL buffer");
    HIP_TRY(hipSetDevice(c->device));
    if (precision == NCE_PREC_FP64) return launch_delta<double, 4>(c, model, occ, R, N, ia, ib, dE, (hipStream_t)stream);
    if (precision == NCE_PREC_MIXED) return launch_delta<float, 4>(c, model, occ, R, N, ia, ib, dE, (hipStream_t)stream);
    return set_err(NCE_ERR_ARG, "unknown precision");
}

// stage a host buffer into device memory, stream-ordered; a ring of pinned buffers guarded by events
static int stage_upload(nce_ctx *c, const void *host, size_t bytes, hipStream_t s, void **dev) {
    nce_ctx::StageRing *ring = nullptr;
    for (nce_ctx::StageRing *g : c->rings)
        if (g->stream == s) ring = g;
    if (!ring) {
        ring = new nce_ctx::StageRing();
        ring->stream = s;
        c->rings.push_back(ring);
        for (auto &st : ring->stage)
            HIP_TRY(hipEventCreateWithFlags(&st.ev, hipEventDisableTiming));
    }
    nce_ctx::Stage &st = ring->stage[ring->next];
    ring->next = (ring->next + 1) % 4;
    if (st.pending) {
        HIP_TRY(hipEventSynchronize(st.ev));
        st.pending = false;
    }
    if (bytes > st.cap) {
        if (st.host) HIP_TRY(hipHostFree(st.host));
        if (st.dev) HIP_TRY(hipFree(st.dev));
        size_t cap = std::max<size_t>(4096, bytes * 2);
        HIP_TRY(hipHostMalloc(&st.host, cap, hipHostMallocDefault));
        HIP_TRY(hipMalloc(&st.dev, cap));
        st.cap = cap;
    }
    std::memcpy(st.host, host, bytes);
    HIP_TRY(hipMemcpyAsync(st.dev, st.host, bytes, hipMemcpyHostToDevice, s));
    HIP_TRY(hipEventRecord(st.ev, s));
    st.pending = true;
    *dev = st.dev;
    return NCE_OK;
}

// A chunk launch decided once: the job table and work items (one staged buffer), the kernel instance, its block
// size and LDS bytes, and whether the moment merge follows.  Shared by nce_pt_run_chunk and the captured cycles.
typedef void (*ChunkKernel)(const ModelDev *, const nce_pt_job *, const int2 *, int64_t);

struct ChunkPlan {
    std::vector<unsigned char> table;                 // [n_jobs] nce_pt_job (16-aligned) then [n_items] int2
    size_t jb_bytes = 0;
    int n_jobs = 0, n_items = 0, block = 0;
    size_t lds = 0;
    int64_t n_steps = 0;
    bool inj = false, merge = false;
    ChunkKernel kernel = nullptr;
};

template <typename Real, int W>
static ChunkKernel general_kernel() { return ce_chunk_kernel<Real, W>; }
template <typename Real, int W>
static ChunkKernel spec_kernel(bool inj) {
    return inj ? ce_chunk_spec_kernel<Real, W, true> : ce_chunk_spec_kernel<Real, W, false>;
}

static int plan_chunk(nce_ctx *c, const nce_pt_job *jobs, int32_t n_jobs, int64_t n_steps, int32_t precision,
                      int32_t tile, ChunkPlan &P) {
    if (!c || !jobs || n_jobs < 1) return set_err(NCE_ERR_ARG, "NULL context / jobs");
    if (n_steps < 0) return set_err(NCE_ERR_ARG, "negative n_steps");
    if (precision != NCE_PREC_FP64 && precision != NCE_PREC_MIXED) return set_err(NCE_ERR_ARG, "unknown precision");
    std::vector<int2> items;
    size_t lds_need = 0;
    bool merge = false;
    for (int j = 0; j < n_jobs; ++j) {
        const nce_pt_job &jb = jobs[j];
        int rc = check_model(c, jb.model);
        if (rc) return rc;
        const ModelDev &m = c->models[jb.model];
        if (jb.n_sites != m.n_sites) return set_err(NCE_ERR_ARG, "job n_sites does not match its model");
        if (jb.n_replicas < 1) return set_err(NCE_ERR_ARG, "job without replicas");
        if (!jb.occ || !jb.energy || !jb.beta || !jb.slot_of_rep || !jb.step || !jb.n_accept || !jb.e_min)
            return set_err(NCE_ERR_ARG, "job has a NULL state buffer");
        if (jb.traj && (jb.sample_interval < 1 || n_steps % jb.sample_interval != 0))
            return set_err(NCE_ERR_ARG, "n_steps must be a multiple of sample_interval");
        if (jb.accumulate && (!jb.e_ref || !jb.s1 || !jb.s2)) return set_err(NCE_ERR_ARG, "accumulators missing");
        if ((jb.acc_n != nullptr) != (jb.acc_mean != nullptr) || (jb.acc_n != nullptr) != (jb.acc_m2 != nullptr))
            return set_err(NCE_ERR_ARG, "slot moments acc_n / acc_mean / acc_m2 must be all set or all NULL");
        if (jb.acc_n && !jb.accumulate) return set_err(NCE_ERR_ARG, "slot moments need accumulate = 1");
        if ((jb.inj_ia != nullptr) != (jb.inj_ib != nullptr) || (jb.inj_ia != nullptr) != (jb.inj_u != nullptr))
            return set_err(NCE_ERR_ARG, "injected streams must be all set or all NULL");
        merge = merge || jb.acc_n != nullptr;
        for (int r = 0; r < jb.n_replicas; ++r) items.push_back(make_int2(j, r));
        lds_need = std::max(lds_need, precision == NCE_PREC_FP64 ? c->lds_chunk64[jb.model] : c->lds_chunk32[jb.model]);
    }
    if (lds_need > 160 * 1024) return set_err(NCE_ERR_UNSUPPORTED, "template does not fit in LDS");
    if (tile != 0 && tile != 1 && tile != 2 && tile != 4 && tile != 8 && tile != 16)
        return set_err(NCE_ERR_ARG, "tile must be 0, 1, 2, 4, 8 or 16");
    const int per_cu = (int)((items.size() + c->n_cu - 1) / c->n_cu);
    // general kernel: waves cooperating on one step of a replica; speculative kernel: steps evaluated in parallel
    int W = tile == 0 ? (per_cu >= 4 ? 1 : per_cu >= 2 ? 2 : 4) : (tile > 8 ? 8 : tile);
    // default speculative tile: 8 waves when several replicas share a CU (16 at one), fewer when the windows of a
    // group's swaps would overlap often (expected pairs W(W-1)/2 * 4 * window / N of the smallest cell above 2: each
    // overlap with an accepted earlier swap costs a re-evaluation).  Measured on cells of 2,048 to 1,048,576 sites
    // at 1 to 12 replicas per CU, 8 waves beat 4 wherever the overlap bound allows 8.
    int tile_spec = (tile == 4 || tile == 8 || tile == 16) ? tile : (per_cu >= 2 ? 8 : 16);
    if (tile == 0) {
        double dens = 0.0;                             // largest window / sites over the jobs
        for (int j = 0; j < n_jobs; ++j) {
            const ModelDev &m = c->models[jobs[j].model];
            dens = std::max(dens, (double)m.Wmax / (double)std::max<int64_t>(m.n_sites, 1));
        }
        while (tile_spec > 4 && 0.5 * tile_spec * (tile_spec - 1) * 4.0 * dens > 2.0) tile_spec /= 2;
    }
    P.jb_bytes = align16(sizeof(nce_pt_job) * n_jobs);
    P.table.assign(P.jb_bytes + items.size() * sizeof(int2), 0);
    std::memcpy(P.table.data(), jobs, sizeof(nce_pt_job) * n_jobs);
    std::memcpy(P.table.data() + P.jb_bytes, items.data(), items.size() * sizeof(int2));
    P.n_jobs = n_jobs;
    P.n_items = (int)items.size();
    P.n_steps = n_steps;
    P.merge = merge;
    // recompute the LDS bytes for the chosen wave count (reduction scratch depends on it)
    size_t lds = 0;
    for (int j = 0; j < n_jobs; ++j) {
        const ModelDev &m = c->models[jobs[j].model];
        lds = std::max(lds, precision == NCE_PREC_FP64 ? chunk_lds_layout<double>(m.nB, m.J, m.Wmax, m.Wcap, W).total
                                                       : chunk_lds_layout<float>(m.nB, m.J, m.Wmax, m.Wcap, W).total);
    }
    bool fast = true;
    const bool inj = jobs[0].inj_ia != nullptr;
    int SW = tile_spec;
    auto spec_lds = [&](int waves) {
        size_t need = 0;
        for (int j = 0; j < n_jobs; ++j) {
            const ModelDev &m = c->models[jobs[j].model];
            need = std::max(need, precision == NCE_PREC_FP64
                                      ? spec_lds_layout<double>(m.J, m.Wcap, waves, m.fs_nlo, m.fs_nfull, m.fs_ntail).total
                                      : spec_lds_layout<float>(m.J, m.Wcap, waves, m.fs_nlo, m.fs_nfull, m.fs_ntail).total);
        }
        return need;
    };
    for (int j = 0; j < n_jobs; ++j) {
        const ModelDev &m = c->models[jobs[j].model];
        if (!m.fs_ok || m.n_sites >= ((int64_t)1 << 31)) fast = false;
        if ((jobs[j].inj_ia != nullptr) != inj) fast = false;          // one stream kind per launch
    }
    if (SW == 16 && spec_lds(16) > 160 * 1024) SW = 8;
    const size_t lds_fast = spec_lds(SW);
    const char *fg = std::getenv("NCE_FORCE_GENERAL");               // tests: run the general kernel on any model
    if (fg && fg[0] == '1') fast = false;
    P.inj = inj;
    if (fast && lds_fast <= 160 * 1024) {
        const bool d = precision == NCE_PREC_FP64;
        P.kernel = SW == 4 ? (d ? spec_kernel<double, 4>(inj) : spec_kernel<float, 4>(inj))
                 : SW == 16 ? (d ? spec_kernel<double, 16>(inj) : spec_kernel<float, 16>(inj))
                            : (d ? spec_kernel<double, 8>(inj) : spec_kernel<float, 8>(inj));
        P.block = 64 * SW;
        P.lds = lds_fast;
    } else {
        const bool d = precision == NCE_PREC_FP64;
        P.kernel = W == 1 ? (d ? general_kernel<double, 1>() : general_kernel<float, 1>())
                 : W == 2 ? (d ? general_kernel<double, 2>() : general_kernel<float, 2>())
                 : W == 4 ? (d ? general_kernel<double, 4>() : general_kernel<float, 4>())
                          : (d ? general_kernel<double, 8>() : general_kernel<float, 8>());
        P.block = 64 * (W > 8 ? 8 : W);
        P.lds = lds;
    }
    if (P.lds > 160 * 1024) return set_err(NCE_ERR_UNSUPPORTED, "template does not fit in LDS");
    return allow_lds(P.kernel, P.lds);
}

// the chunk kernel of a plan on its device tables, then the moment merge when a job asks for it
static int launch_plan(const ChunkPlan &P, const ModelDev *d_models, const void *d_table, hipStream_t s) {
    const nce_pt_job *d_jobs = (const nce_pt_job *)d_table;
    const int2 *d_items = (const int2 *)((const unsigned char *)d_table + P.jb_bytes);
    int64_t n_steps = P.n_steps;
    void *args[] = {(void *)&d_models, (void *)&d_jobs, (void *)&d_items, (void *)&n_steps};
    HIP_TRY(hipLaunchKernel((const void *)P.kernel, dim3(P.n_items), dim3(P.block), args, P.lds, s));
    if (P.merge) {
        const int nb = (P.n_items + MERGE_BS - 1) / MERGE_BS;
        hipLaunchKernelGGL(pt_merge_kernel, dim3(nb), dim3(MERGE_BS), 0, s, d_jobs, d_items, P.n_items, P.n_steps);
        HIP_TRY(hipGetLastError());
    }
    return NCE_OK;
}

extern "C" int nce_pt_run_chunk(nce_ctx *c, const nce_pt_job *jobs, int32_t n_jobs, int64_t n_steps,
                                int32_t precision, int32_t tile, void *stream) {
    if (!c || !jobs || n_jobs < 1) return set_err(NCE_ERR_ARG, "NULL context / jobs");
    if (n_steps == 0) return NCE_OK;
    HIP_TRY(hipSetDevice(c->device));
    ChunkPlan P;
    if (int rc = plan_chunk(c, jobs, n_jobs, n_steps, precision, tile, P)) return rc;
    hipStream_t s = (hipStream_t)stream;
    void *dev = nullptr;
    if (int rc = stage_upload(c, P.table.data(), P.table.size(), s, &dev)) return rc;
    return launch_plan(P, c->d_models, dev, s);
}

extern "C" int nce_pt_exchange(nce_ctx *c, const nce_pt_exchange_job *jobs, int32_t n_jobs, void *stream) {
    if (!c || !jobs || n_jobs < 1) return set_err(NCE_ERR_ARG, "NULL context / jobs");
    for (int j = 0; j < n_jobs; ++j) {
        const nce_pt_exchange_job &e = jobs[j];
        if (e.n_replicas < 1 || !e.energy || !e.rep_of_slot || !e.slot_of_rep || !e.beta_slot || !e.beta_rep ||
            !e.n_swaps)
            return set_err(NCE_ERR_ARG, "exchange job has a NULL buffer");
    }
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t s = (hipStream_t)stream;
    void *dev = nullptr;
    int rc = stage_upload(c, jobs, sizeof(nce_pt_exchange_job) * n_jobs, s, &dev);
    if (rc) return rc;
    hipLaunchKernelGGL(pt_exchange_kernel, dim3(n_jobs), dim3(EXCH_BS), 0, s, (const nce_pt_exchange_job *)dev);
    HIP_TRY(hipGetLastError());
    return NCE_OK;
}

// ─────────────────────────────────── captured exchange cycles (HIP graphs) ───────────────────────────────────
struct nce_pt_cycle {
    nce_ctx *c = nullptr;
    void *d_mem = nullptr;            // model descriptors (a snapshot: the context's array may be reallocated),
                                      // job table + items, exchange table
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" void nce_pt_cycle_destroy(nce_pt_cycle *cy) {
    if (!cy) return;
    if (cy->c) (void)hipSetDevice(cy->c->device);
    if (cy->exec) (void)hipGraphExecDestroy(cy->exec);
    if (cy->graph) (void)hipGraphDestroy(cy->graph);
    if (cy->d_mem) {
        (void)hipDeviceSynchronize();                 // a replay still in flight reads the tables
        (void)hipFree(cy->d_mem);
    }
    delete cy;
}

extern "C" int nce_pt_cycle_create(nce_ctx *c, const nce_pt_job *jobs, int32_t n_jobs, int64_t n_steps,
                                   int32_t chunks_per_cycle, int32_t precision, int32_t tile,
                                   const nce_pt_exchange_job *ex_jobs, int32_t n_ex_jobs, nce_pt_cycle **out) {
    if (!out) return set_err(NCE_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (!c || !jobs || n_jobs < 1) return set_err(NCE_ERR_ARG, "NULL context / jobs");
    if (n_steps < 1 || chunks_per_cycle < 1) return set_err(NCE_ERR_ARG, "a cycle needs n_steps >= 1 and chunks >= 1");
    if (n_ex_jobs < 0 || (n_ex_jobs > 0 && !ex_jobs)) return set_err(NCE_ERR_ARG, "bad exchange table");
    for (int j = 0; j < n_jobs; ++j)
        if (jobs[j].inj_ia) return set_err(NCE_ERR_ARG, "injected streams cannot be replayed by a cycle");
    for (int j = 0; j < n_ex_jobs; ++j) {
        const nce_pt_exchange_job &e = ex_jobs[j];
        if (e.n_replicas < 1 || !e.energy || !e.rep_of_slot || !e.slot_of_rep || !e.beta_slot || !e.beta_rep ||
            !e.n_swaps)
            return set_err(NCE_ERR_ARG, "exchange job has a NULL buffer");
        if (!e.ex_count) return set_err(NCE_ERR_ARG, "a cycle's exchange jobs must give ex_count");
        if (e.inj_u || e.inj_parity >= 0) return set_err(NCE_ERR_ARG, "injected exchanges cannot be replayed");
    }
    HIP_TRY(hipSetDevice(c->device));
    ChunkPlan P;
    if (int rc = plan_chunk(c, jobs, n_jobs, n_steps, precision, tile, P)) return rc;
    const size_t models_bytes = align16(c->models.size() * sizeof(ModelDev));
    const size_t table_bytes = align16(P.table.size());
    const size_t ex_bytes = sizeof(nce_pt_exchange_job) * (size_t)n_ex_jobs;
    nce_pt_cycle *cy = new nce_pt_cycle();
    cy->c = c;
    auto fail = [&](int rc) {
        nce_pt_cycle_destroy(cy);
        return rc;
    };
    if (hipMalloc(&cy->d_mem, models_bytes + table_bytes + ex_bytes) != hipSuccess)
        return fail(set_err(NCE_ERR_NOMEM, "cycle tables: hipMalloc failed"));
    unsigned char *base = (unsigned char *)cy->d_mem;
    std::vector<unsigned char> host(models_bytes + table_bytes + ex_bytes, 0);
    std::memcpy(host.data(), c->models.data(), c->models.size() * sizeof(ModelDev));
    std::memcpy(host.data() + models_bytes, P.table.data(), P.table.size());
    if (n_ex_jobs) std::memcpy(host.data() + models_bytes + table_bytes, ex_jobs, ex_bytes);
    if (hipMemcpy(cy->d_mem, host.data(), host.size(), hipMemcpyHostToDevice) != hipSuccess)
        return fail(set_err(NCE_ERR_HIP, "cycle tables: hipMemcpy failed"));
    hipStream_t cap = nullptr;
    if (hipStreamCreateWithFlags(&cap, hipStreamNonBlocking) != hipSuccess)
        return fail(set_err(NCE_ERR_HIP, "cycle: hipStreamCreateWithFlags failed"));
    int rc = NCE_OK;
    if (hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        rc = set_err(NCE_ERR_HIP, "cycle: hipStreamBeginCapture failed");
    } else {
        for (int k = 0; k < chunks_per_cycle && rc == NCE_OK; ++k)
            rc = launch_plan(P, (const ModelDev *)base, base + models_bytes, cap);
        if (rc == NCE_OK && n_ex_jobs) {
            hipLaunchKernelGGL(pt_exchange_kernel, dim3(n_ex_jobs), dim3(EXCH_BS), 0, cap,
                               (const nce_pt_exchange_job *)(base + models_bytes + table_bytes));
            if (hipGetLastError() != hipSuccess) rc = set_err(NCE_ERR_HIP, "cycle: exchange launch failed");
        }
        const hipError_t e = hipStreamEndCapture(cap, &cy->graph);      // always end the capture
        if (rc == NCE_OK && e != hipSuccess) rc = set_err(NCE_ERR_HIP, "cycle: hipStreamEndCapture failed");
    }
    (void)hipStreamDestroy(cap);
    if (rc != NCE_OK) return fail(rc);
    if (hipGraphInstantiate(&cy->exec, cy->graph, nullptr, nullptr, 0) != hipSuccess)
        return fail(set_err(NCE_ERR_HIP, "cycle: hipGraphInstantiate failed"));
    *out = cy;
    return NCE_OK;
}

extern "C" int nce_pt_cycle_launch(nce_pt_cycle *cy, int32_t n_cycles, void *stream) {
    if (!cy || !cy->exec) return set_err(NCE_ERR_ARG, "NULL cycle");
    if (n_cycles < 0) return set_err(NCE_ERR_ARG, "negative cycle count");
    HIP_TRY(hipSetDevice(cy->c->device));
    for (int k = 0; k < n_cycles; ++k) HIP_TRY(hipGraphLaunch(cy->exec, (hipStream_t)stream));
    return NCE_OK;
}

#include "nn_update.hip"
