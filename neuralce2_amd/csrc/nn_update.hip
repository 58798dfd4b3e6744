// Neural cluster expansion (NeuralCE) kernels of libneuralce_hip.so, included into neuralce_hip.hip.
//
// Model (neuralce/nce/local_update_implicit.py:43-112, neuralce/nce/general.py:213-235): h_0 = standardised layer 0
// (h0s); hidden layer l = 1..L at site s
//   pre_l(s)  = h_{l-1}(s) W_self + sum_p mean_{q in shell p of s} h_{l-1}(q) W_p + b,
//   gpre_l(s) = sum_p mean_{q in shell p of s} h_{l-1}(q) G_p + gb                  (gated models only),
//   h_l(s)    = tanh(pre_l) * sigmoid(gpre_l)  (gate)  |  tanh(pre_l)  |  pre_l * sigmoid(pre_l)  (SiLU);
// site energy e(s) = h0s(s) . lin + h_L(s) . w; E = sum_s e / nA + b (mean readout) or sum_s e + b.
//
// Per-replica cache (elements of Real): h0s [N][c0] (cache mode 0 only) | for each hidden layer: pre [N][d]
// (| gpre [N][d]).
//
// Kernels:
//   nn_h0_kernel        standardised layer 0 of every site from the layer-0 window and reverse entries
//   nn_layer_kernel     pre-activations of one hidden layer at every site
//   nn_energy_kernel    site energies and fixed-order partial sums
//   nn_chunk_kernel     swap-Metropolis steps of one replica per workgroup with the receptive-field local update
//                       in incremental form.  pre_l is linear in h_{l-1}: its change on ball_l is the change of
//                       h_{l-1} on ball_{l-1} projected through [W_self | W_p / n_p | G_p / n_p] and added at every
//                       shell partner; the new activations follow from the cached pre-activations plus the change,
//                       and the energy change is the layer-0 change times lin plus the last layer's change times w.
//                       Per step: the balls of the two flips are built in an LDS hash table (two-phase claim /
//                       assign, no spin-waits), the layer-0 change is scattered into the ball-0 rows with LDS
//                       atomics, the partner hits of every layer are sorted by target (counting sort in LDS) so the
//                       sums are gathers, and on acceptance the changes are added to the caches.

template <typename Real>
__device__ inline const Real *nn_w(const NNModelDev &m) {
    if constexpr (sizeof(Real) == 8) return m.wts;
    else return m.wts_f;
}
template <typename Real>
__device__ inline const Real *nn_lin(const NNModelDev &m) {
    if constexpr (sizeof(Real) == 8) return m.lin;
    else return m.lin_f;
}
template <typename Real>
__device__ inline const Real *nn_headw(const NNModelDev &m) {
    if constexpr (sizeof(Real) == 8) return m.head_w;
    else return m.head_w_f;
}
// Activations.  FP64: the C library functions.  FP32: from the hardware base-2 exponential and reciprocal
// (v_exp_f32, v_rcp_f32, about 1 ulp each): sigmoid(x) = 1 / (1 + 2^(-x log2 e)), tanh(x) = 1 - 2 / (1 + 2^(2x log2 e));
// the absolute error stays at a few 1e-8, the FP32 rounding level of the values themselves.
__device__ inline double nn_sigmoid(double x) { return 1.0 / (1.0 + exp(-x)); }
__device__ inline double nn_tanh(double x) { return tanh(x); }
__device__ inline float nn_sigmoid(float x) {
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.4426950408889634f * x));
}
__device__ inline float nn_tanh(float x) {
    return 1.0f - 2.0f * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(2.8853900817779268f * x));
}

template <typename Real>
__device__ inline Real nn_act(const NNModelDev &m, Real pre, Real gpre) {
    if (m.act == 1) return pre * nn_sigmoid(pre);
    Real out = nn_tanh(pre);
    if (m.has_gate) out *= nn_sigmoid(gpre);
    return out;
}
// the same with the kind fixed at compile time: 0 tanh, 1 SiLU, 2 gated tanh
template <int ACT, typename Real>
__device__ inline Real nn_act_k(Real pre, Real gpre) {
    if constexpr (ACT == 1) return pre * nn_sigmoid(pre);
    else if constexpr (ACT == 2) return nn_tanh(pre) * nn_sigmoid(gpre);
    else return nn_tanh(pre);
}

__device__ inline int64_t nn_site(const NNModelDev &m, int64_t center, int4 e) {
    const uint32_t c = (uint32_t)center;
    const uint32_t d12 = (uint32_t)(m.d1 * m.d2);
    const int k0c = (int)udiv_magic(c, m.div12_m, m.div12_s, d12);
    const int rem = (int)c - k0c * (int)d12;
    const int k1c = (int)udiv_magic((uint32_t)rem, m.div2_m, m.div2_s, (uint32_t)m.d2);
    const int k2c = rem - k1c * m.d2;
    int k0 = k0c + e.y, k1 = k1c + e.z, k2 = k2c + e.w;
    k0 -= (k0 >= m.d0) ? m.d0 : 0;
    k1 -= (k1 >= m.d1) ? m.d1 : 0;
    k2 -= (k2 >= m.d2) ? m.d2 : 0;
    return ((int64_t)k0 * m.d1 + k1) * m.d2 + k2;
}

// the cell vector ib - ia outside ball_L - ball_L: the two flips' balls share no site
__device__ inline bool nn_balls_apart(const NNModelDev &m, int64_t ia, int64_t ib) {
    const uint32_t d12 = (uint32_t)(m.d1 * m.d2);
    int k[2][3];
    for (int f = 0; f < 2; ++f) {
        const uint32_t c = (uint32_t)(f == 0 ? ia : ib);
        const int k0 = (int)udiv_magic(c, m.div12_m, m.div12_s, d12);
        const int rem = (int)c - k0 * (int)d12;
        const int k1 = (int)udiv_magic((uint32_t)rem, m.div2_m, m.div2_s, (uint32_t)m.d2);
        k[f][0] = k0; k[f][1] = k1; k[f][2] = rem - k1 * m.d2;
    }
    int v0 = k[1][0] - k[0][0], v1 = k[1][1] - k[0][1], v2 = k[1][2] - k[0][2];
    v0 += v0 < 0 ? m.d0 : 0;
    v1 += v1 < 0 ? m.d1 : 0;
    v2 += v2 < 0 ? m.d2 : 0;
    const uint32_t v = (uint32_t)((v0 * m.d1 + v1) * m.d2 + v2);
    return ((gptr(m.pre_near)[v >> 5] >> (v & 31)) & 1u) == 0;
}

// ── full cache ─────────────────────────────────────────────────────────────────────────────────────────────
constexpr int NN_H0_SITES = 64;

// LDS bytes per site of the window-sign rows: an odd number of 4-byte words, so the rows start in different banks;
// 0 (members read from memory) where the rows and the channel accumulators would pass 96 KB per workgroup
__host__ __device__ inline int nn_h0_sign_stride(int W0, int C0, int real_bytes) {
    const int stride = 4 * (((W0 + 3) / 4) | 1);
    return (size_t)NN_H0_SITES * ((size_t)C0 * real_bytes + stride) <= 96 * 1024 ? stride : 0;
}

template <typename Real>
__global__ __launch_bounds__(NN_H0_SITES) void nn_h0_kernel(const NNModelDev *__restrict__ models, int model,
                                                            const int8_t *__restrict__ occ_all, int64_t N,
                                                            Real *__restrict__ h0_all, int64_t h0_stride) {
    extern __shared__ __align__(16) unsigned char smem[];
    const NNModelDev &m = models[model];
    Real *acc = (Real *)smem;                          // [NN_H0_SITES][C0]
    const int r = blockIdx.y;
    const int64_t site = (int64_t)blockIdx.x * NN_H0_SITES + threadIdx.x;
    const int8_t *occ = occ_all + (int64_t)r * N;
    Real *row = acc + (size_t)threadIdx.x * m.C0;
    for (int c = 0; c < m.C0; ++c) row[c] = (Real)0;
    // the spin (+1 / -1) of every window slot of this thread's site, read once: [NN_H0_SITES][stride] bytes
    const int stride = nn_h0_sign_stride(m.W0, m.C0, (int)sizeof(Real));
    int8_t *sg = (int8_t *)(acc + (size_t)NN_H0_SITES * m.C0) + (size_t)threadIdx.x * stride;
    if (site < N) {
        if (stride > 0)
            for (int k = 0; k < m.W0; ++k) sg[k] = occ[nn_site(m, site, m.win[k])] ? 1 : -1;
        for (int e = 0; e < m.J0; ++e) {
            const uint2 sl = m.ent_slot[e];
            const uint32_t s4[4] = {sl.x & 0xffffu, sl.x >> 16, sl.y & 0xffffu, sl.y >> 16};
            double prod = 1.0;
            for (int q = 0; q < 4; ++q) {
                if (s4[q] == 0xffffu) continue;
                if (stride > 0) {
                    prod *= (double)sg[s4[q]];
                } else {
                    const int64_t x = nn_site(m, site, m.win[s4[q]]);
                    prod *= occ[x] ? 1.0 : -1.0;
                }
            }
            row[m.ent_col[e]] += (Real)(prod * m.ent_coef[e]);
        }
        row[0] = (Real)1;
        Real *out = h0_all + (int64_t)r * h0_stride + site * m.C0;
        for (int c = 0; c < m.C0; ++c) out[c] = (Real)(((double)row[c] - m.mu[c]) / m.sd[c]);
    }
}

constexpr int NN_LDC = 32;               // output channels per thread in the full recompute of a layer

// pre-activations of hidden layer l at every site (inputs: h0s for l = 0, else the activations of layer l - 1).
// One site and NN_LDC output channels per thread (the channel block is the same for the whole workgroup, so the
// weights are read as broadcasts): the previous layer's activations and the shell means are formed once per block
// of channels, and every output channel sums its terms in the order of the layer formula.
template <typename Real>
__global__ __launch_bounds__(256) void nn_layer_kernel(const NNModelDev *__restrict__ models, int model, int l,
                                                       int64_t N, Real *__restrict__ cache_all,
                                                       const Real *__restrict__ h0_all, int64_t h0_stride) {
    const NNModelDev &m = models[model];
    const int r = blockIdx.y;
    const int D = m.dd[l], C = m.cin[l];
    const int nch = (D + NN_LDC - 1) / NN_LDC;
    const int d0 = (int)(blockIdx.x % nch) * NN_LDC;
    const int64_t site = (int64_t)(blockIdx.x / nch) * 256 + threadIdx.x;
    if (site >= N) return;
    Real *cache = cache_all + (int64_t)r * m.cache_elems;
    const Real *W = nn_w<Real>(m);
    const Real *ip = l == 0 ? h0_all + (int64_t)r * h0_stride : cache + m.off_pre[l - 1];
    const Real *ig = (l > 0 && m.has_gate) ? cache + m.off_gpre[l - 1] : nullptr;
    auto get = [&](int64_t x, int c) -> Real {
        if (l == 0) return ip[x * C + c];
        return nn_act<Real>(m, ip[x * C + c], ig ? ig[x * C + c] : (Real)0);
    };
    Real pre[NN_LDC], gpre[NN_LDC];
#pragma unroll
    for (int j = 0; j < NN_LDC; ++j) {
        const int d = d0 + j;
        pre[j] = d < D ? W[m.off_b[l] + d] : (Real)0;
        gpre[j] = m.has_gate && d < D ? W[m.off_gb[l] + d] : (Real)0;
    }
    for (int c = 0; c < C; ++c) {
        const Real hs = get(site, c);
        const Real *ws = W + m.off_self[l] + (int64_t)c * D + d0;
#pragma unroll
        for (int j = 0; j < NN_LDC; ++j)
            if (d0 + j < D) pre[j] += hs * ws[j];
    }
    int q0 = 0;
    for (int p = 0; p < m.P; ++p) {
        const int np_ = m.shell_count[p];
        const Real inv = np_ > 0 ? (Real)1 / (Real)np_ : (Real)0;
        for (int c = 0; c < C; ++c) {
            Real z = (Real)0;
            for (int q = 0; q < np_; ++q) z += get(nn_site(m, site, m.part[q0 + q]), c);
            z *= inv;
            const Real *wsh = W + m.off_shell[l] + ((int64_t)p * C + c) * D + d0;
            const Real *wg = W + m.off_gate[l] + ((int64_t)p * C + c) * D + d0;
#pragma unroll
            for (int j = 0; j < NN_LDC; ++j) {
                if (d0 + j >= D) continue;
                pre[j] += z * wsh[j];
                if (m.has_gate) gpre[j] += z * wg[j];
            }
        }
        q0 += np_;
    }
#pragma unroll
    for (int j = 0; j < NN_LDC; ++j) {
        if (d0 + j >= D) continue;
        cache[m.off_pre[l] + site * D + d0 + j] = pre[j];
        if (m.has_gate) cache[m.off_gpre[l] + site * D + d0 + j] = gpre[j];
    }
}

// site energies and per-block partial sums in fixed order
template <typename Real>
__global__ __launch_bounds__(256) void nn_energy_kernel(const NNModelDev *__restrict__ models, int model, int64_t N,
                                                        const Real *__restrict__ cache_all,
                                                        const Real *__restrict__ h0_all, int64_t h0_stride,
                                                        double *__restrict__ partial) {
    __shared__ double red[4];
    const NNModelDev &m = models[model];
    const int r = blockIdx.y;
    const int64_t site = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const Real *cache = cache_all + (int64_t)r * m.cache_elems;
    double e = 0.0;
    if (site < N) {
        const Real *h0 = h0_all + (int64_t)r * h0_stride + site * m.C0;
        const Real *lin = nn_lin<Real>(m);
        Real acc = (Real)0;
        for (int c = 0; c < m.C0; ++c) acc += h0[c] * lin[c];
        if (m.L > 0) {
            const int l = m.L - 1, D = m.dd[l];
            const Real *pre = cache + m.off_pre[l] + site * D;
            const Real *gp = m.has_gate ? cache + m.off_gpre[l] + site * D : nullptr;
            const Real *hw = nn_headw<Real>(m);
            for (int d = 0; d < D; ++d) acc += nn_act<Real>(m, pre[d], gp ? gp[d] : (Real)0) * hw[d];
        }
        e = (double)acc;
    }
    e = wave_sum(e);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
    __syncthreads();
    if (threadIdx.x == 0) partial[(int64_t)r * gridDim.x + blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ void nn_energy_finish_kernel(const NNModelDev *__restrict__ models, int model,
                                        const double *__restrict__ partial, int n_part, double *E) {
    const NNModelDev &m = models[model];
    const int r = blockIdx.x;
    double acc = 0.0;
    for (int i = threadIdx.x; i < n_part; i += 64) acc += partial[(int64_t)r * n_part + i];
    acc = wave_sum(acc);
    if (threadIdx.x == 0) E[r] = (m.readout == 0 ? acc / (double)m.nA : acc) + m.head_b;
}

// ── local update chunk kernel ─────────────────────────────────────────────────────────────────────────────
constexpr int NN_BS = 1024;                // threads per replica: 16 waves, 4 per SIMD
constexpr int NN_EMPTY = -1;
constexpr int NN_RB = 4;                 // rows per thread in the projection
constexpr int NN_EU = 4;                 // layer-0 entries per thread in flight together
constexpr int NN_CU = 4;                 // cache rows per thread in flight together at a commit
constexpr int NN_LA = 8;                 // proposals drawn, and their species read, ahead of the steps (power of 2)
// one drawn proposal and the species of its two sites at the step it runs
struct NNProp {
    double u;
    int ia, ib, si, sj;
};

struct NNLds {
    size_t keys, vals, slot_site, cnt, red, misc, la, wv, part, w, dh0, list, pr, total;
};
__host__ __device__ inline NNLds nn_lds_layout(int hash_cap, int ball_cap, int win_n, int n_part, int n_c0, size_t w_bytes,
                                               size_t dh0_bytes, size_t list_bytes, size_t pr_bytes,
                                               size_t val_bytes = sizeof(int), bool with_cnt = true) {
    NNLds L;
    size_t o = 0;
    L.keys = o;      o = align16(o + (size_t)hash_cap * sizeof(int));
    L.vals = o;      o = align16(o + (size_t)hash_cap * val_bytes);
    L.slot_site = o; o = align16(o + (size_t)ball_cap * sizeof(int));
    L.cnt = o;       o = align16(o + (with_cnt ? (size_t)(ball_cap + 1) * sizeof(int) : 0));
    L.red = o;       o = align16(o + (NN_BS / 64 + 1) * sizeof(double));
    L.misc = o;      o = align16(o + (4 + NN_MAXL + 1) * sizeof(int));
    L.la = o;        o = align16(o + NN_LA * sizeof(NNProp));
    L.wv = o;        o = align16(o + (size_t)win_n * 3 * sizeof(int));
    L.part = o;      o = align16(o + (size_t)n_part * (sizeof(int4) + 2 * sizeof(int)) + (size_t)n_c0 * sizeof(int));
    L.w = o;         o = align16(o + w_bytes);
    L.dh0 = o;       o = align16(o + dh0_bytes);
    L.list = o;      o = align16(o + list_bytes);
    L.pr = o;        o = align16(o + pr_bytes);
    L.total = o;
    return L;
}
// fused: the BF16 layers, whose stacked matrix is staged instead of the projection matrix and which keep neither a
// hit list nor a projection block in LDS; their hash table (hash_cap_f) holds uint16 slots and no hit counts
template <typename Real>
__host__ __device__ inline NNLds nn_lds_for(const NNModelDev &m, int mode, bool fused = false) {
    return nn_lds_layout(fused ? m.hash_cap_f : m.hash_cap, m.ball_cap[m.L], 2 * m.W0, m.n_part, m.c0_eff,
                         fused ? (size_t)m.wf_bytes : (size_t)m.w_elems * sizeof(Real),
                         mode >= 1 ? (size_t)m.ball_cap[0] * m.C0 * sizeof(Real) : 0,
                         mode >= 2 && !fused ? (size_t)m.hit_elems * sizeof(int) : 0,
                         mode >= 3 && !fused ? (size_t)m.pr_elems * sizeof(Real) : 0,
                         fused ? sizeof(uint16_t) : sizeof(int), !fused);
}

// multiplicative hashing: the position is the top log2(cap) bits of the 32-bit product (cap = mask + 1 >= 64, a power
// of two). The low bits would depend on site mod cap alone, so sites a multiple of cap apart (one cell layer along the
// slowest axis, for instance) would always share a position and lengthen the probe chains.
__device__ inline uint32_t nn_hash(int site, int mask) {
    return (((uint32_t)site * 2654435761u) >> __clz(mask)) & (uint32_t)mask;
}

// phase 1 of an insertion: claim or find the key; returns the table position and whether this thread claimed it
__device__ inline int nn_hash_claim(int *keys, int mask, int site, bool *claimed) {
    uint32_t h = nn_hash(site, mask);
    while (true) {
        const int prev = atomicCAS(&keys[h], NN_EMPTY, site);
        if (prev == NN_EMPTY) { *claimed = true; return (int)h; }
        if (prev == site) { *claimed = false; return (int)h; }
        h = (h + 1) & (uint32_t)mask;
    }
}

template <typename V>
__device__ inline int nn_hash_find(const int *keys, const V *vals, int mask, int site) {
    uint32_t h = nn_hash(site, mask);
    while (true) {
        const int k = keys[h];
        if (k == site) return (int)vals[h];
        if (k == NN_EMPTY) return -1;
        h = (h + 1) & (uint32_t)mask;
    }
}

// block-wide sum, the same in every thread.  red is read after the function returns, up to the caller's next
// barrier: nothing may write red before that barrier (the chunk kernel's next write is steps of barriers later)
template <typename Real>
__device__ inline Real nn_block_sum(Real v, Real *red) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    Real s = red[0];
    for (int i = 1; i < NN_BS / 64; ++i) s += red[i];
    return s;
}

// in-place exclusive prefix sum of a[0..n) (block-wide); returns the total
__device__ inline int nn_block_scan(int *a, int n, double *red_d) {
    int *red = (int *)red_d;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int seg = (n + NN_BS - 1) / NN_BS;
    const int b = tid * seg, e = min(b + seg, n);
    int loc = 0;
    for (int i = b; i < e; ++i) loc += a[i];
    int incl = loc;
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o, 64);
        if (lane >= o) incl += v;
    }
    if (lane == 63) red[w] = incl;
    __syncthreads();
    int wave_off = 0, total = 0;
    for (int i = 0; i < NN_BS / 64; ++i) {
        if (i < w) wave_off += red[i];
        total += red[i];
    }
    int run = wave_off + incl - loc;
    for (int i = b; i < e; ++i) {
        const int v = a[i];
        a[i] = run;
        run += v;
    }
    __syncthreads();
    return total;
}

// four consecutive channels (16-byte aligned rows: every cache / scratch region starts at a multiple of 4 elements)
template <typename Real>
struct alignas(16) NNv4 {
    Real v[4];
};
// four channels at p, moved as one vector access in p's address space (LDS or global)
template <typename Real>
__device__ inline NNv4<Real> nn_ld4(const Real *p) {
    typedef Real vec __attribute__((ext_vector_type(4)));
    const vec x = *(const vec *)p;
    return NNv4<Real>{{x[0], x[1], x[2], x[3]}};
}
template <typename Real>
__device__ inline NNv4<Real> nn_ld4(const NCE_GLOBAL Real *p) {
    typedef Real vec __attribute__((ext_vector_type(4)));
    const vec x = *(const NCE_GLOBAL vec *)p;
    return NNv4<Real>{{x[0], x[1], x[2], x[3]}};
}
template <typename Real>
struct alignas(2 * sizeof(Real)) NNv2 {
    Real v[2];
};
template <typename Real>
__device__ inline NNv2<Real> nn_ld2(const Real *p) {
    typedef Real vec __attribute__((ext_vector_type(2)));
    const vec x = *(const vec *)p;
    return NNv2<Real>{{x[0], x[1]}};
}
template <typename Real>
__device__ inline NNv2<Real> nn_ld2(const NCE_GLOBAL Real *p) {
    typedef Real vec __attribute__((ext_vector_type(2)));
    const vec x = *(const NCE_GLOBAL vec *)p;
    return NNv2<Real>{{x[0], x[1]}};
}
template <typename Real>
__device__ inline void nn_st2(NCE_GLOBAL Real *p, const NNv2<Real> &v) {
    typedef Real vec __attribute__((ext_vector_type(2)));
    *(NCE_GLOBAL vec *)p = vec{v.v[0], v.v[1]};
}
template <typename Real>
__device__ inline void nn_st4(NCE_GLOBAL Real *p, const NNv4<Real> &v) {
    typedef Real vec __attribute__((ext_vector_type(4)));
    *(NCE_GLOBAL vec *)p = vec{v.v[0], v.v[1], v.v[2], v.v[3]};
}
typedef __bf16 nn_bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 nn_bf16x4 __attribute__((ext_vector_type(4)));
typedef float nn_floatx4 __attribute__((ext_vector_type(4)));
constexpr int NN_FCT = 4;                // BF16 layers: output tiles of 16 columns per pass over the hits
constexpr int NN_FCH = 4;                // BF16 layers: groups of four neighbour-table columns loaded together
constexpr int NN_FMAXT = 8;              // BF16 layers: at most 8 output tiles (D (1 + gate) <= 128)

// A replica over S workgroups (BF16 layers, cooperative launch of n_items * S workgroups, n_items a multiple of 8):
// workgroup b runs part (b / 8) % S of item (b / 8 / S) * 8 + b % 8, so the S parts of an item share one XCD (the
// dispatcher places workgroup b on XCD b % 8; each workgroup checks its XCC id) and its L2.  They exchange rows
// through that L2: a group barrier drains the stores, counts on a per-item counter (relaxed atomics at agent scope,
// performed in the L2) and, after its bounded wait, invalidates the CU's L1.  A wait that runs out, or a workgroup
// on another XCD, sets the abort flag: no later barrier waits, and the replica's energy is returned as NaN.
struct NNSplit {
    int32_t S;                   // parts per item (1: one workgroup per replica)
    uint32_t *bar;               // [n_items] barrier counters, zero at launch
    double *part;                // [n_items][S] energy partials of the current step
};
constexpr uint32_t NN_SPLIT_SPIN = 1u << 22;       // polls of a barrier counter before the wait gives up (~0.1 s)

template <typename Real, int MODE, bool MFMA, bool SPLIT = false>
__global__ __launch_bounds__(NN_BS) void nn_chunk_kernel(const NNModelDev *__restrict__ models,
                                                         const nce_nn_job *__restrict__ jobs,
                                                         const int2 *__restrict__ items, int64_t n_steps,
                                                         double *__restrict__ eval_dE, const NNSplit sp) {
    extern __shared__ __align__(16) unsigned char smem[];
    int item = blockIdx.x, prt = 0;
    if constexpr (SPLIT) {
        const int local = blockIdx.x >> 3;
        prt = local % sp.S;
        item = (local / sp.S) * 8 + (blockIdx.x & 7);
    }
    const int2 it = items[item];
    const nce_nn_job job = jobs[it.x];
    const NNModelDev &m = models[job.model];
    const int r = it.y;
    const int tid = threadIdx.x;
    const int L = m.L, C0 = m.C0, n_part = m.n_part, P = m.P;
    // hash table: the BF16 layers' one is larger (shorter probe chains) and holds uint16 slots
    const int hcap = MFMA ? m.hash_cap_f : m.hash_cap;
    const int mask = hcap - 1;
    const NNLds lay = nn_lds_for<Real>(m, MODE, MFMA);
    using HV = typename std::conditional<MFMA, uint16_t, int>::type;
    int *keys = (int *)(smem + lay.keys);
    HV *vals = (HV *)(smem + lay.vals);
    int *slot_site = (int *)(smem + lay.slot_site);
    int *cnt = (int *)(smem + lay.cnt);
    double *red = (double *)(smem + lay.red);
    int *misc = (int *)(smem + lay.misc);
    // the two layer-0 windows of a step: site, hash position (then slot) and species of every window slot (w_sp: bit 0
    // the species before the slot's flip, bit 1 after it, the slot above them; the second window sees the first flip
    // applied)
    int *w_site = (int *)(smem + lay.wv), *w_pos = w_site + 2 * m.W0, *w_sp = w_pos + 2 * m.W0;
    // shell partners (offset and shell of each) and the used layer-0 channels, staged once per launch
    int4 *s_part = (int4 *)(smem + lay.part);
    int *s_part_shell = (int *)(s_part + m.n_part);
    int *s_part_col = s_part_shell + m.n_part;
    int *s_c0 = s_part_col + m.n_part;
    for (int k = tid; k < m.n_part; k += NN_BS) {
        s_part[k] = m.part[k];
        s_part_shell[k] = m.part_shell[k];
        s_part_col[k] = m.part_col[k];
    }
    for (int k = tid; k < m.c0_eff; k += NN_BS) s_c0[k] = m.c0_ch[k];
    for (int i = tid; i < hcap; i += NN_BS) keys[i] = NN_EMPTY;
    if (tid == 0) {
        misc[0] = 0;
        misc[1] = 0;                                   // abort flag of a split replica
        if constexpr (SPLIT)
            if (__builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) != (blockIdx.x & 7u)) misc[1] = 1;   // XCC_ID
    }
    if constexpr (MFMA) {
        // the zero row of both input-row buffers of the BF16 layers (read for absent hits and self terms)
        NCE_GLOBAL __bf16 *xb = (NCE_GLOBAL __bf16 *)(gptr((Real *)job.scratch) + (int64_t)r * m.scratch_elems +
                                                      m.soff_pr);
        const int n8 = m.xb_stride >> 3;
        if (tid < 2 * n8)
            *(NCE_GLOBAL nn_bf16x8 *)(xb + (tid < n8 ? 0 : m.xb_off1) + (int64_t)m.xb_zrow * m.xb_stride +
                                      8 * (tid < n8 ? tid : tid - n8)) = nn_bf16x8{};
    }
    __syncthreads();

    // buffers in device memory are addressed as global (flat accesses would also count against lgkmcnt, so every
    // LDS wait would drain the cache and scratch traffic in flight)
    NCE_GLOBAL int8_t *occ = gptr(job.occ) + (int64_t)r * job.n_sites;
    NCE_GLOBAL Real *cache = gptr((Real *)job.cache) + (int64_t)r * m.cache_elems;
    NCE_GLOBAL Real *scr = gptr((Real *)job.scratch) + (int64_t)r * m.scratch_elems;
    Real *wl = (Real *)(smem + lay.w);                 // projection matrix of the current layer
    // per-step arrays: in LDS where MODE places them, else in the replica's scratch
    std::conditional_t<(MODE >= 1), Real *, NCE_GLOBAL Real *> dh0;
    std::conditional_t<(MODE >= 2), int *, NCE_GLOBAL int *> list;
    std::conditional_t<(MODE >= 3), Real *, NCE_GLOBAL Real *> pr;
    if constexpr (MODE >= 1) dh0 = (Real *)(smem + lay.dh0); else dh0 = scr + m.soff_dh0;
    if constexpr (MODE >= 2) list = (int *)(smem + lay.list); else list = (NCE_GLOBAL int *)(scr + m.soff_list);
    if constexpr (MODE >= 3) pr = (Real *)(smem + lay.pr); else pr = scr + m.soff_pr;
    NCE_GLOBAL int *tgt = (NCE_GLOBAL int *)(scr + m.soff_tgt);
    const NCE_GLOBAL Real *W = gptr(nn_w<Real>(m));
    const NCE_GLOBAL Real *lin = gptr(nn_lin<Real>(m));
    const NCE_GLOBAL Real *hw = gptr(nn_headw<Real>(m));
    const NCE_GLOBAL int *win = (const NCE_GLOBAL int *)gptr(m.win);
    const NCE_GLOBAL uint64_t *ent_slot = (const NCE_GLOBAL uint64_t *)gptr(m.ent_slot);
    const NCE_GLOBAL int *ent_col = gptr(m.ent_col);
    const NCE_GLOBAL double *ent_q = gptr(m.ent_q);
    double E = job.energy[r];
    const double beta = job.beta[r];
    const uint64_t step0 = (uint64_t)job.step[r];
    int64_t nacc = 0;
    const int64_t inj_base = (int64_t)r * n_steps;
    int *counts = misc + 4;                            // slots of ball_0 .. ball_L this step (same in every thread)
    bool w_staged = false;                             // the projection matrix in LDS is the current layer's
    bool pre_in_list = false;                          // list and cnt hold the fixed table (one-layer models)
    // group barrier of the S parts of a split replica (a workgroup barrier when S = 1)
    uint32_t nbar = 0;
    auto gbar = [&]() {
        if constexpr (SPLIT) {
            __builtin_amdgcn_s_waitcnt(0);             // this wave's stores have reached the L2
            __syncthreads();
            ++nbar;
            if (tid == 0 && !misc[1]) {
                uint32_t *c = sp.bar + item;
                __hip_atomic_fetch_add(c, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const uint32_t want = nbar * (uint32_t)sp.S;
                uint32_t polls = 0;
                while (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < want) {
                    __builtin_amdgcn_s_sleep(2);
                    if (++polls > NN_SPLIT_SPIN) {
                        misc[1] = 1;
                        break;
                    }
                }
            }
            __syncthreads();
            asm volatile("buffer_inv sc0" ::: "memory");   // the other parts' rows are read from the L2
        } else {
            __syncthreads();
        }
    };
    const int S = SPLIT ? sp.S : 1;
    const bool builder = prt == 0;                     // part that builds a hash step's balls and tables
    auto n_own = [&](int n) { return (n - prt + S - 1) / S; };   // of rows 0 .. n-1, those of this part: prt + S k

    NNProp *s_la = (NNProp *)(smem + lay.la);
    for (int64_t t = 0; t < n_steps; ++t) {
        const int la_slot = (int)(t & (NN_LA - 1));
        if (la_slot == 0) {
            // the next NN_LA proposals, one lane each, with the species of their sites read together (a swap
            // accepted before a proposal's step patches them below)
            if (tid < NN_LA && t + tid < n_steps) {
                int64_t pa, pb;
                double pu;
                if (job.inj_ia) {
                    pa = gptr(job.inj_ia)[inj_base + t + tid];
                    pb = gptr(job.inj_ib)[inj_base + t + tid];
                    pu = gptr(job.inj_u)[inj_base + t + tid];
                } else {
                    const Proposal p =
                        draw_proposal_1b(job.seed, step0 + (uint64_t)(t + tid), (uint32_t)r, (uint64_t)m.n_cells);
                    pa = p.ia; pb = p.ib; pu = p.u;
                }
                s_la[tid] = NNProp{pu, (int)pa, (int)pb, (int)occ[pa], (int)occ[pb]};
            }
            __syncthreads();
        }
        const NNProp pq = s_la[la_slot];
        const int64_t ia = pq.ia, ib = pq.ib;
        const double u = pq.u;
        const int si = pq.si, sj = pq.sj;
        if (ia == ib || si == sj) {                      // no-op proposal: dE = 0, accepted (reference semantics)
            if (eval_dE) {
                if (tid == 0) eval_dE[blockIdx.x] = 0.0;
            } else {
                ++nacc;
            }
        } else {
            // balls of two flips that share no site: every slot, hit and target range comes from the fixed tables
            const bool fixed = m.pre_ok && nn_balls_apart(m, ia, ib);
            // ball 0: the two layer-0 windows (w_pos: the hash position of each window slot, then its slot)
            for (int k = tid; k < 2 * m.W0; k += NN_BS) {
                const NCE_GLOBAL int *wk = win + 4 * (k < m.W0 ? k : k - m.W0);
                const int site = (int)nn_site(m, k < m.W0 ? ia : ib, make_int4(0, wk[1], wk[2], wk[3]));
                if (fixed) {
                    w_pos[k] = gptr(m.pre_win)[k];
                } else if (builder) {
                    bool claimed;
                    const int pos = nn_hash_claim(keys, mask, site, &claimed);
                    if (claimed) {
                        const int s = atomicAdd(&misc[0], 1);
                        vals[pos] = (HV)s;
                        slot_site[s] = site;
                    }
                    w_pos[k] = pos;
                }
                w_site[k] = site;
                const bool f2 = k >= m.W0;
                int sp = occ[site];
                if (f2 && site == ia) sp = sj;                // flip 2 sees flip 1 applied
                int spn = sp;
                if (site == (f2 ? ib : ia)) spn = f2 ? si : sj;
                w_sp[k] = (sp ? 1 : 0) | (spn ? 2 : 0) | (fixed ? w_pos[k] << 2 : 0);
            }
            if (fixed) {
                const NCE_GLOBAL int *ps = (const NCE_GLOBAL int *)gptr(m.pre_slot);
                for (int s = tid; s < 2 * m.pre_n[L]; s += NN_BS) {
                    const NCE_GLOBAL int *e = ps + 4 * s;
                    slot_site[s] = (int)nn_site(m, e[0] ? ib : ia, make_int4(0, e[1], e[2], e[3]));
                }
            }
            __syncthreads();
            if constexpr (MFMA) {
                // BF16 layers: every ball and every layer's neighbour table before the layer-0 change
                if (fixed) {
                    for (int l = 0; l <= L; ++l) counts[l] = 2 * m.pre_n[l];
                } else if (builder) {
                    counts[0] = misc[0];
                    for (int k = tid; k < 2 * m.W0; k += NN_BS) {
                        w_pos[k] = vals[w_pos[k]];
                        w_sp[k] |= w_pos[k] << 2;
                    }
                    // hit i = (q, kk) = (i / n_part, i % n_part), stepped without divisions
                    const int q_0 = n_part > 0 ? tid / n_part : 0, kk_0 = tid - q_0 * n_part;
                    const int dq = n_part > 0 ? NN_BS / n_part : 0, dk = NN_BS - dq * n_part;
                    for (int l = 0; l < L; ++l) {
                        // ball_{l+1}: the partners of ball_l's sites, claimed (new sites get the next slots) for
                        // the sites new in ball_l only (an older site's partners are in ball_l already)
                        const int q_lo = l == 0 ? 0 : counts[l - 1], q_hi = counts[l];
                        int q = q_lo + q_0, kk = kk_0;
                        for (int i = q_lo * n_part + tid; i < q_hi * n_part; i += NN_BS, q += dq, kk += dk) {
                            if (kk >= n_part) { kk -= n_part; ++q; }
                            const int site = (int)nn_site(m, slot_site[q], s_part[kk]);
                            bool claimed;
                            const int pos = nn_hash_claim(keys, mask, site, &claimed);
                            if (claimed) {
                                const int sl = atomicAdd(&misc[0], 1);
                                vals[pos] = (HV)sl;
                                slot_site[sl] = site;
                            }
                        }
                        __syncthreads();
                        counts[l + 1] = misc[0];
                    }
                    // one neighbour table for every layer: row t of ball_L, the source slot of each partner's hit
                    // (a target and a partner have at most one source), the zero row where none; layer l reads the
                    // sources below counts[l] (the slots are nested) and takes any other as the zero row
                    NCE_GLOBAL uint32_t *nb32 = (NCE_GLOBAL uint32_t *)(scr + m.soff_nbr);
                    const uint32_t z2 = (uint32_t)m.xb_zrow * 0x10001u;
                    for (int e = tid; e < (counts[L] * m.nbr_w) >> 1; e += NN_BS) nb32[e] = z2;
                    __syncthreads();
                    NCE_GLOBAL uint16_t *nbt = (NCE_GLOBAL uint16_t *)(scr + m.soff_nbr);
                    // (each hit's partner found again in the hash, every slot now final: LDS reads only)
                    int q = q_0, kk = kk_0;
                    for (int i = tid; i < counts[L - 1] * n_part; i += NN_BS, q += dq, kk += dk) {
                        if (kk >= n_part) { kk -= n_part; ++q; }
                        const int t = nn_hash_find(keys, vals, mask, (int)nn_site(m, slot_site[q], s_part[kk]));
                        if (t >= 0) nbt[(int64_t)t * m.nbr_w + s_part_col[kk]] = (uint16_t)q;
                    }
                    __syncthreads();
                }
                if (SPLIT && !fixed) {
                    // the builder publishes the slots (counts, window slots, sites); the other parts copy them
                    NCE_GLOBAL int *pub = (NCE_GLOBAL int *)(scr + m.soff_pub);
                    const int npub = L + 1 + 2 * m.W0 + counts[L];
                    if (builder)
                        for (int k = tid; k < npub; k += NN_BS)
                            pub[k] = k <= L ? counts[k] : k < L + 1 + 2 * m.W0 ? w_pos[k - L - 1]
                                                                               : slot_site[k - L - 1 - 2 * m.W0];
                    gbar();
                    if (!builder) {
                        // (clamped into range, so that rows of an aborted step still address memory in bounds)
                        for (int k = tid; k <= L; k += NN_BS) counts[k] = max(0, min(pub[k], m.ball_cap[k]));
                        __syncthreads();
                        for (int k = tid; k < 2 * m.W0; k += NN_BS) {
                            w_pos[k] = max(0, min(pub[L + 1 + k], counts[0] - 1));
                            w_sp[k] |= w_pos[k] << 2;
                        }
                        for (int k = tid; k < counts[L]; k += NN_BS)
                            slot_site[k] = max(0, min(pub[L + 1 + 2 * m.W0 + k], (int)m.n_sites - 1));
                    }
                }
                __syncthreads();
                for (int k = tid; k < counts[0] * C0; k += NN_BS) dh0[k] = (Real)0;
            } else {
            counts[0] = fixed ? 2 * m.pre_n[0] : misc[0];
            for (int k = tid; k < counts[0] * C0; k += NN_BS) dh0[k] = (Real)0;
            if (!fixed)
                for (int k = tid; k < 2 * m.W0; k += NN_BS) {
                    w_pos[k] = vals[w_pos[k]];
                    w_sp[k] |= w_pos[k] << 2;
                }
            }
            __syncthreads();
            // layer-0 change of flip 1, then of flip 2 with flip 1 applied.  A round loads NN_EU entries per thread
            // (slots, coefficient, channel) before it uses any, and both flips use them.
            for (int e0 = tid; e0 < m.J0; e0 += NN_EU * NN_BS) {
                uint64_t esl[NN_EU];
                double eq[NN_EU];
                int ecol[NN_EU];
#pragma unroll
                for (int u = 0; u < NN_EU; ++u) {
                    const int e = min(e0 + u * NN_BS, m.J0 - 1);
                    esl[u] = ent_slot[e];
                    eq[u] = ent_q[e];
                    ecol[u] = ent_col[e];
                }
#pragma unroll
                for (int u = 0; u < NN_EU; ++u) {
                    if (e0 + u * NN_BS >= m.J0) break;
                    const uint32_t lo = (uint32_t)esl[u], hi = (uint32_t)(esl[u] >> 32);
                    const uint32_t s4[4] = {lo & 0xffffu, lo >> 16, hi & 0xffffu, hi >> 16};
                    for (int f = 0; f < 2; ++f) {
                        const int wb = f == 0 ? 0 : m.W0;      // this flip's window in w_pos / w_sp
                        int xs[4];
                        double po = 1.0, pn = 1.0;
                        for (int q = 0; q < 4; ++q) {
                            if (s4[q] == 0xffffu) { xs[q] = -1; continue; }
                            xs[q] = wb + (int)s4[q];
                            const int b = w_sp[xs[q]];         // species before and after this flip
                            po *= (b & 1) ? 1.0 : -1.0;
                            pn *= (b & 2) ? 1.0 : -1.0;
                        }
                        // (pn - po) is 0 or +-2, so (pn - po) * (coef / sd) rounds as ((pn - po) * coef) / sd does
                        const double dv = (pn - po) * eq[u];
                        if (dv != 0.0) {
                            const Real v = (Real)dv;
                            for (int q = 0; q < 4; ++q) {
                                if (xs[q] < 0) continue;
                                const int s = w_sp[xs[q]] >> 2;   // the member's slot
                                if ((s & (S - 1)) != prt) continue;   // a split replica's part keeps its own rows
                                atomicAdd((Real *)&dh0[(int64_t)s * C0 + ecol[u]], v);
                            }
                        }
                    }
                }
            }
            __syncthreads();
            double part = 0.0;
            {
                // the layer-0 term over this part's rows (all rows unless the replica is split); element k of those
                // rows is channel k % C0, stepped without divisions
                const int dcs = NN_BS % C0;
                int c = tid % C0;
                for (int k = tid; k < n_own(counts[0]) * C0; k += NN_BS) {
                    const int so = S == 1 ? 0 : k / C0;
                    part += (double)(dh0[(int64_t)(S == 1 ? k : (prt + so * S) * C0 + c)] * lin[c]);
                    c += dcs;
                    c -= c >= C0 ? C0 : 0;
                }
            }
            if constexpr (MFMA) {
                // the first layer's input rows: the used channels of the layer-0 change as bf16, zero up to cpad[0]
                const int n8 = m.cpad[0] >> 3;
                NCE_GLOBAL __bf16 *x0 = (NCE_GLOBAL __bf16 *)(scr + m.soff_pr);
                for (int k = tid; k < n_own(counts[0]) * n8; k += NN_BS) {   // (a split replica: alternate rows)
                    const int so = k / n8, c0 = (k - so * n8) * 8, s = prt + so * S;
                    nn_bf16x8 v;
#pragma unroll
                    for (int i = 0; i < 8; ++i) {
                        const int c = c0 + i;
                        v[i] = (__bf16)(c < m.c0_eff ? (float)dh0[(int64_t)s * C0 + s_c0[c]] : 0.0f);
                    }
                    *(NCE_GLOBAL nn_bf16x8 *)(x0 + (int64_t)s * m.xb_stride + c0) = v;
                }
                gbar();                                // every part's input rows are in place
            }
            // hidden layers: projection of the previous layer's change, partner hits sorted by target, gather
            for (int l = 0; l < L; ++l) {
                const int Cin = l == 0 ? m.c0_eff : m.cin[l], D = m.dd[l], M = m.mcat[l];
                const int prev = counts[l];
                // the previous layer's change: layer 0 in dh0 (used channels only), deeper layers in the scratch
                const NCE_GLOBAL Real *dxg = scr + m.soff_dh[(l - 1) & 1];
                auto dx = [&](int64_t row, int c) -> Real {
                    return l == 0 ? (Real)dh0[row * C0 + s_c0[c]] : dxg[row * m.cin[l] + c];
                };
                const NCE_GLOBAL Real *Wc = W + m.off_wcat[l];
                // (1) pr[q][j] = dx[q] . Wcat[:, j]
                if constexpr (MFMA) {
                    // BF16 layers: the stacked matrix (layout at nn_stack_bf16), built at model upload and copied
                    // 16 bytes per thread
                    if (!w_staged) {
                        const int n16 = (D * (m.has_gate ? 2 : 1) * (m.kf[l] + 8)) >> 3;
                        typedef int i4 __attribute__((ext_vector_type(4)));
                        const NCE_GLOBAL i4 *src = (const NCE_GLOBAL i4 *)(gptr(m.wf_img) + m.wf_off[l]);
                        for (int k = tid; k < n16; k += NN_BS) ((i4 *)wl)[k] = src[k];
                    }
                } else if (sizeof(Real) == 4 && m.mfma_ok) {
                    // FP32 MFMA (v_mfma_f32_16x16x4_f32): the layer's matrix staged in LDS as [c][j]; one 16 x 16
                    // output tile per wave at a time, CT column tiles per pass.  Lane l holds A[row l%16][k l/16],
                    // B[k l/16][col l%16], C[rows 4(l/16)..+3][col l%16].
                    if (!w_staged) {
                        for (int k = tid; k < Cin * M; k += NN_BS) {
                            const int c = k / M, j = k - c * M;
                            wl[k] = Wc[(int64_t)(l == 0 ? s_c0[c] : c) * M + j];
                        }
                        __syncthreads();
                    }
                    const int lane = tid & 63, wv = tid >> 6;
                    const int nrt = (prev + 15) >> 4, nct = M >> 4;
                    constexpr int CT = 4;
                    for (int task = wv; task < nrt * ((nct + CT - 1) / CT); task += NN_BS / 64) {
                        const int rt = task % nrt, cg = (task / nrt) * CT;
                        const int row = rt * 16 + (lane & 15), kq = lane >> 4;
                        nn_floatx4 acc[CT];
#pragma unroll
                        for (int t = 0; t < CT; ++t) acc[t] = nn_floatx4{0.0f, 0.0f, 0.0f, 0.0f};
                        if (sizeof(Real) == 4 && l > 0 && (Cin & 15) == 0) {
                            // input rows in the scratch: lane group kq takes the k range kq * Cin/4 .. +Cin/4 - 1, so
                            // its next four k are consecutive channels of its row (one 16-byte load); the sum over k
                            // is the same, only its order differs
                            const int Kq = Cin >> 2;
                            const NCE_GLOBAL Real *rp = dxg + (int64_t)row * Cin + kq * Kq;
                            for (int t0 = 0; t0 < Kq; t0 += 4) {
                                NNv4<Real> av = NNv4<Real>{{(Real)0, (Real)0, (Real)0, (Real)0}};
                                if (row < prev) av = nn_ld4<Real>(rp + t0);
#pragma unroll 1
                                for (int u = 0; u < 4; ++u) {
                                    const int c = kq * Kq + t0 + u;
                                    const float a = (float)(u == 0 ? av.v[0] : u == 1 ? av.v[1] : u == 2 ? av.v[2] : av.v[3]);
#pragma unroll
                                    for (int t = 0; t < CT; ++t) {
                                        if (cg + t < nct) {
                                            const float b = (float)wl[c * M + (cg + t) * 16 + (lane & 15)];
                                            acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[t], 0, 0, 0);
                                        }
                                    }
                                }
                            }
                        } else {
                        for (int k0 = 0; k0 < Cin; k0 += 4) {
                            const int c = k0 + kq;
                            float a = 0.0f;
                            if (row < prev && c < Cin) a = (float)dx(row, c);
#pragma unroll
                            for (int t = 0; t < CT; ++t) {
                                if (cg + t < nct) {
                                    const float b = c < Cin ? (float)wl[c * M + (cg + t) * 16 + (lane & 15)] : 0.0f;
                                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[t], 0, 0, 0);
                                }
                            }
                        }
                        }
#pragma unroll
                        for (int t = 0; t < CT; ++t) {
                            if (cg + t >= nct) continue;
#pragma unroll
                            for (int i = 0; i < 4; ++i) {
                                const int r0 = rt * 16 + 4 * (lane >> 4) + i;
                                if (r0 < prev) pr[(int64_t)r0 * M + (cg + t) * 16 + (lane & 15)] = (Real)acc[t][i];
                            }
                        }
                    }
                } else {
                // the layer's matrix (used rows only) staged in LDS, 4 x 4 register tiles
                if (!w_staged) {
                    for (int k = tid; k < Cin * M; k += NN_BS) {
                        const int c = k / M, j = k - c * M;
                        wl[k] = Wc[(int64_t)(l == 0 ? s_c0[c] : c) * M + j];
                    }
                    __syncthreads();
                }
                const int nqb = (prev + NN_RB - 1) / NN_RB, ncb = (M + 3) / 4;
                for (int k = tid; k < nqb * ncb; k += NN_BS) {
                    const int qb = k / ncb, j0 = (k - qb * ncb) * 4;
                    const int q0 = qb * NN_RB;
                    Real acc[NN_RB][4];
#pragma unroll
                    for (int i = 0; i < NN_RB; ++i)
#pragma unroll
                        for (int u = 0; u < 4; ++u) acc[i][u] = (Real)0;
                    if (l > 0 && (Cin & 1) == 0) {
                        // input rows in the scratch: two consecutive channels of each row per load
                        typedef Real vec2 __attribute__((ext_vector_type(2)));
                        for (int c = 0; c < Cin; c += 2) {
                            Real wa[4], wb[4], xa[NN_RB], xb[NN_RB];
#pragma unroll
                            for (int i = 0; i < NN_RB; ++i) {
                                vec2 x = vec2{(Real)0, (Real)0};
                                if (q0 + i < prev) x = *(const NCE_GLOBAL vec2 *)(dxg + (int64_t)(q0 + i) * Cin + c);
                                xa[i] = x[0];
                                xb[i] = x[1];
                            }
#pragma unroll
                            for (int u = 0; u < 4; ++u) {
                                wa[u] = j0 + u < M ? wl[c * M + j0 + u] : (Real)0;
                                wb[u] = j0 + u < M ? wl[(c + 1) * M + j0 + u] : (Real)0;
                            }
#pragma unroll
                            for (int i = 0; i < NN_RB; ++i)
#pragma unroll
                                for (int u = 0; u < 4; ++u) acc[i][u] += xa[i] * wa[u] + xb[i] * wb[u];
                        }
                    } else {
                    for (int c = 0; c < Cin; ++c) {
                        Real wv[4], xv[NN_RB];
#pragma unroll
                        for (int u = 0; u < 4; ++u) wv[u] = j0 + u < M ? wl[c * M + j0 + u] : (Real)0;
#pragma unroll
                        for (int i = 0; i < NN_RB; ++i) xv[i] = q0 + i < prev ? dx(q0 + i, c) : (Real)0;
#pragma unroll
                        for (int i = 0; i < NN_RB; ++i)
#pragma unroll
                            for (int u = 0; u < 4; ++u) acc[i][u] += xv[i] * wv[u];
                    }
                    }
#pragma unroll
                    for (int i = 0; i < NN_RB; ++i)
#pragma unroll
                        for (int u = 0; u < 4; ++u)
                            if (q0 + i < prev && j0 + u < M) pr[(int64_t)(q0 + i) * M + j0 + u] = acc[i][u];
                }
                }
                // (2) partner sites of ball_l, (3) the hits sorted by target: list holds (source slot << 8 | shell)
                // grouped by target, cnt[s] the end of target s's range (the BF16 layers built their balls and
                // neighbour tables before the layer-0 change)
                const int nh = prev * n_part;
                int nb;
                if constexpr (MFMA) {
                    nb = counts[l + 1];
                    __syncthreads();                     // the stacked matrix is staged
                } else if (fixed) {
                    nb = 2 * m.pre_n[l + 1];
                    counts[l + 1] = nb;
                    if (!pre_in_list) {
                        const NCE_GLOBAL int *hl = gptr(m.pre_list) + m.pre_list_off[l];
                        const NCE_GLOBAL int *he = gptr(m.pre_end) + m.pre_end_off[l];
                        for (int i = tid; i < nh; i += NN_BS) list[i] = hl[i];
                        for (int s = tid; s < nb; s += NN_BS) cnt[s] = he[s];
                    }
                    pre_in_list = L == 1;                // one layer: the table stays in place until a hash step
                    __syncthreads();
                } else {
                pre_in_list = false;
                // claim (new sites get the next slots), then the slot of every hit
                // hit i = (q, kk) = (i / n_part, i % n_part), stepped without divisions
                const int q_0 = n_part > 0 ? tid / n_part : 0, kk_0 = tid - q_0 * n_part;
                const int dq = n_part > 0 ? NN_BS / n_part : 0, dk = NN_BS - dq * n_part;
                int q = q_0, kk = kk_0;
                for (int i = tid; i < nh; i += NN_BS, q += dq, kk += dk) {
                    if (kk >= n_part) { kk -= n_part; ++q; }
                    const int site = (int)nn_site(m, slot_site[q], s_part[kk]);
                    bool claimed;
                    const int pos = nn_hash_claim(keys, mask, site, &claimed);
                    if (claimed) {
                        const int s = atomicAdd(&misc[0], 1);
                        vals[pos] = (HV)s;
                        slot_site[s] = site;
                    }
                    tgt[i] = pos;
                }
                __syncthreads();
                nb = misc[0];
                counts[l + 1] = nb;
                for (int s = tid; s < nb; s += NN_BS) cnt[s] = 0;
                __syncthreads();
                for (int i = tid; i < nh; i += NN_BS) {
                    const int s = vals[tgt[i]];
                    tgt[i] = s;
                    atomicAdd(&cnt[s], 1);
                }
                __syncthreads();
                // counting sort of the hits by target: after the fill cnt[s] is the end of target s's range
                nn_block_scan(cnt, nb, red);
                q = q_0; kk = kk_0;
                for (int i = tid; i < nh; i += NN_BS, q += dq, kk += dk) {
                    if (kk >= n_part) { kk -= n_part; ++q; }
                    const int pos = atomicAdd(&cnt[tgt[i]], 1);
                    list[pos] = (q << 8) | s_part_shell[kk];
                }
                __syncthreads();
                }
                // (4) change of the pre-activations, new activations, change of the layer output
                // the targets' new pre-activations (cached value + change), copied into the cache at a commit
                NCE_GLOBAL Real *npre = scr + m.soff_npre[l];
                NCE_GLOBAL Real *ngpre = scr + m.soff_ngpre[l];
                NCE_GLOBAL Real *dh = scr + m.soff_dh[l & 1];
                const NCE_GLOBAL Real *cpre = cache + m.off_pre[l];
                const NCE_GLOBAL Real *cgpre = cache + m.off_gpre[l];
                const bool last = l + 1 == L;
                const int og = D * (1 + P);
                constexpr int U = 4;                     // items per thread in flight together
                if constexpr (MFMA) {
                // BF16 layers, one wave per 16 targets (lane: target tile*16 + lane%16, channel group lane/16):
                // the change of the pre-activations of target t is C = Wstack^T . [x_t | sum_p hi | sum_p lo], a
                // 16-column tile of MFMA v_mfma_f32_16x16x32_bf16 per 16 output columns.  The B operand is built
                // in registers: x_t is the input row (bf16, the zero row for t outside ball_{l-1}); sum_p is the
                // FP32 sum of the input rows of t's hits in shell p, split into bf16 high and low parts (their sum
                // equals it to 2^-16 relative).  A lane holds C[columns 4 g .. + 3 of each tile][target], so the
                // cache rows, the scratch rows and the next layer's input rows move four channels at a time.
                // The hits of a target come from its neighbour-table row (fixed width, the zero row where a
                // partner lies outside ball_{l-1}), so every lane loads the same number of rows, with no branches.
                const int Ks = m.kf[l] + 8;
                const int nct = (D * (m.has_gate ? 2 : 1)) >> 4;
                const __bf16 *wt = (const __bf16 *)wl;
                const uint32_t xrow = 2u * (uint32_t)m.xb_stride;   // bytes per input row
                const int zrow = m.xb_zrow;
                const NCE_GLOBAL char *xin =
                    (const NCE_GLOBAL char *)((const NCE_GLOBAL __bf16 *)(scr + m.soff_pr) + (l & 1) * m.xb_off1);
                NCE_GLOBAL __bf16 *xout = (NCE_GLOBAL __bf16 *)(scr + m.soff_pr) + ((l + 1) & 1) * m.xb_off1;
                const NCE_GLOBAL uint16_t *nbt = fixed ? gptr(m.pre_nbr) + m.pre_nbr_off[l]
                                                       : (const NCE_GLOBAL uint16_t *)(scr + m.soff_nbr);
                const int nbw = m.nbr_w, nps = m.nbr_ps;
                const int lane = tid & 63, g = lane >> 4, tl = lane & 15;
                const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
                // (tiles of NN_FCT output tiles each: a gated model of D = 64 passes over the hits twice)
                const int ngrp = (nct + NN_FCT - 1) / NN_FCT;
                const int ntask = ((nb + 15) >> 4) * ngrp;
                // NSB 32-wide k steps per segment (input width 32 NSB): in step s lane group g takes channels
                // 32 s + 8 g .. +7, so the four lane groups of a target read 64 contiguous bytes of each row
                auto layer = [&]<int NSB>() {
                    constexpr int NZ = 8 * NSB;
                    const uint32_t cob = 16u * (uint32_t)g;   // byte offset of the lane's channels in step 0
                    // (the parts of a split replica interleave, so a layer's last, partial round of tiles is spread
                    // over all of them)
                    for (int task = wave * S + prt; task < ntask; task += S * (NN_BS / 64)) {
                        const int tile = task / ngrp, c0t = (task - tile * ngrp) * NN_FCT;
                        const int t = tile * 16 + tl;
                        const bool tv = t < nb;
                        // A fragment rows of the group's output tiles (a tile past the last reads the last: its
                        // products are computed and dropped, so no MFMA waits on a branch)
                        const __bf16 *wrow[NN_FCT];
#pragma unroll
                        for (int ct = 0; ct < NN_FCT; ++ct)
                            wrow[ct] = wt + (min(c0t + ct, nct - 1) * 16 + tl) * Ks + 8 * g;
                        nn_floatx4 acc[NN_FCT];
#pragma unroll
                        for (int ct = 0; ct < NN_FCT; ++ct) acc[ct] = nn_floatx4{0.0f, 0.0f, 0.0f, 0.0f};
                        auto mma = [&](int k, const nn_bf16x8 &bv) {
                            nn_bf16x8 av[NN_FCT];
#pragma unroll
                            for (int ct = 0; ct < NN_FCT; ++ct) av[ct] = *(const nn_bf16x8 *)(wrow[ct] + k);
#pragma unroll
                            for (int ct = 0; ct < NN_FCT; ++ct)
                                acc[ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[ct], bv, acc[ct], 0, 0, 0);
                        };
                        auto ldrow = [&](int row, int half) -> nn_bf16x8 {
                            return *(const NCE_GLOBAL nn_bf16x8 *)(xin + (uint32_t)row * xrow + cob + 64u * half);
                        };
                        {
                            const int xr = t < prev ? t : zrow;
#pragma unroll
                            for (int s = 0; s < NSB; ++s) mma(32 * s, ldrow(xr, s));
                        }
                        // the selectors of the bf16 sums live in registers: the compiler encodes the pair (1, 0) as
                        // the inline constant 1.0, which the hardware reads as the 32-bit float 1.0, i.e. (0, 1)
                        typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
                        bf2 lo1 = bf2{(__bf16)1.0f, (__bf16)0.0f}, hi1 = bf2{(__bf16)0.0f, (__bf16)1.0f};
                        asm volatile("" : "+v"(lo1));
                        asm volatile("" : "+v"(hi1));
                        // the target's neighbour-table row (target 0's for a lane past the last target)
                        const NCE_GLOBAL uint16_t *tb = nbt + (int64_t)(tv ? t : 0) * nbw;
                        for (int p = 0; p < P; ++p) {
                            float z[NZ];
#pragma unroll
                            for (int e = 0; e < NZ; ++e) z[e] = 0.0f;
                            // a row's channels added with v_dot2_f32_bf16 against (1, 0) and (0, 1): one instruction
                            // per channel, exact products, one FP32 rounding per addition (equal to the FP32 sum)
                            auto add = [&](const nn_bf16x8 &v, int o) {
#pragma unroll
                                for (int e = 0; e < 4; ++e) {
                                    const bf2 pr2 = bf2{v[2 * e], v[2 * e + 1]};
                                    z[o + 2 * e] = __builtin_amdgcn_fdot2_f32_bf16(pr2, lo1, z[o + 2 * e], false);
                                    z[o + 2 * e + 1] = __builtin_amdgcn_fdot2_f32_bf16(pr2, hi1, z[o + 2 * e + 1], false);
                                }
                            };
                            // shell p's columns in groups of four (one 8-byte table load each): NN_FCH groups' table
                            // entries load together, then each group's four rows (columns without a partner in
                            // ball_{l-1} hold the zero row)
                            typedef uint32_t u2 __attribute__((ext_vector_type(2)));
                            const NCE_GLOBAL u2 *tp = (const NCE_GLOBAL u2 *)(tb + p * nps);
                            const int nch = nps >> 2;
                            for (int c0 = 0; c0 < nch; c0 += NN_FCH) {
                                u2 e4[NN_FCH];
#pragma unroll
                                for (int cc = 0; cc < NN_FCH; ++cc) e4[cc] = tp[min(c0 + cc, nch - 1)];
#pragma unroll
                                for (int cc = 0; cc < NN_FCH; ++cc) {
                                    if (c0 + cc >= nch) break;
                                    // (a hash step's table serves every layer: a source at or past prev, outside
                                    // ball_{l-1}, reads the zero row; the test also keeps the rows of an aborted
                                    // split step in bounds)
                                    auto rsel = [&](uint32_t v) -> int { return (int)v < prev ? (int)v : zrow; };
                                    const int rw[4] = {rsel(e4[cc].x & 0xffffu), rsel(e4[cc].x >> 16),
                                                       rsel(e4[cc].y & 0xffffu), rsel(e4[cc].y >> 16)};
                                    // (a column without a partner adds the zero row: its lanes load nothing)
                                    nn_bf16x8 v[4][NSB];
#pragma unroll
                                    for (int u = 0; u < 4; ++u)
#pragma unroll
                                        for (int s = 0; s < NSB; ++s)
                                            v[u][s] = rw[u] != zrow ? ldrow(rw[u], s) : nn_bf16x8{};
#pragma unroll
                                    for (int u = 0; u < 4; ++u)
#pragma unroll
                                        for (int s = 0; s < NSB; ++s) add(v[u][s], 8 * s);
                                }
                            }
#pragma unroll
                            for (int s = 0; s < NSB; ++s) {
                                nn_bf16x8 hi, lo;
#pragma unroll
                                for (int e = 0; e < 8; ++e) {
                                    const float x = z[8 * s + e];
                                    hi[e] = (__bf16)x;
                                    lo[e] = (__bf16)(x - (float)hi[e]);
                                }
                                const int k = (1 + 2 * p) * 32 * NSB + 32 * s;
                                mma(k, hi);
                                mma(k + 32 * NSB, lo);
                            }
                        }
                        if (tv) {
                            using V4 = NNv4<Real>;
                            const V4 z4 = V4{{(Real)0, (Real)0, (Real)0, (Real)0}};
                            const int64_t so = (int64_t)slot_site[t] * D;
                            // the cached rows of every output quad of the group first: a store to the scratch ahead
                            // of a cache load would keep the load from moving up (the pointers may alias), so each
                            // quad would wait on the memory in turn (a quad past the last reads channels in range)
                            constexpr int NQ = NN_FCT;
                            V4 pq[NQ], gq[NQ];
#pragma unroll
                            for (int q = 0; q < NQ; ++q) {
                                const bool gate = m.has_gate;
                                if (gate && q >= NQ / 2) break;
                                const int jc = min(((gate ? c0t / 2 : c0t) + q) * 16 + 4 * g, D - 4);
                                pq[q] = nn_ld4<Real>(cpre + so + jc);
                                gq[q] = gate ? nn_ld4<Real>(cgpre + so + jc) : z4;
                            }
                            // four output channels jc .. jc + 3 of target t (ACT: the activation, fixed per tile)
                            auto out = [&]<int ACT>(int jc, const nn_floatx4 &ap, const nn_floatx4 *ag, const V4 &pv,
                                                    const V4 &gv) {
                                const V4 dp = V4{{ap[0], ap[1], ap[2], ap[3]}};
                                V4 dg = z4, hv, nv, ng;
                                if (ag) dg = V4{{(*ag)[0], (*ag)[1], (*ag)[2], (*ag)[3]}};
#pragma unroll
                                for (int e = 0; e < 4; ++e) {
                                    nv.v[e] = pv.v[e] + dp.v[e];
                                    ng.v[e] = gv.v[e] + dg.v[e];
                                    hv.v[e] = nn_act_k<ACT, Real>(nv.v[e], ng.v[e]) - nn_act_k<ACT, Real>(pv.v[e], gv.v[e]);
                                }
                                nn_st4<Real>(npre + (int64_t)t * D + jc, nv);
                                if (ag) nn_st4<Real>(ngpre + (int64_t)t * D + jc, ng);
                                if (last) {
#pragma unroll
                                    for (int e = 0; e < 4; ++e) part += (double)(hv.v[e] * hw[jc + e]);
                                } else {
                                    nn_bf16x4 hb;
#pragma unroll
                                    for (int e = 0; e < 4; ++e) hb[e] = (__bf16)(float)hv.v[e];
                                    *(NCE_GLOBAL nn_bf16x4 *)(xout + (int64_t)t * m.xb_stride + jc) = hb;
                                }
                            };
                            if (m.has_gate) {
                                // gated models alternate pre-activation and gate tiles (tanh times the gate's sigmoid)
#pragma unroll
                                for (int ct = 0; ct < NN_FCT / 2; ++ct)
                                    if (c0t + 2 * ct < nct)
                                        out.template operator()<2>((c0t / 2 + ct) * 16 + 4 * g, acc[2 * ct], &acc[2 * ct + 1],
                                                                   pq[ct], gq[ct]);
                            } else if (m.act == 1) {
#pragma unroll
                                for (int ct = 0; ct < NN_FCT; ++ct)
                                    if (c0t + ct < nct)
                                        out.template operator()<1>((c0t + ct) * 16 + 4 * g, acc[ct], nullptr, pq[ct], z4);
                            } else {
#pragma unroll
                                for (int ct = 0; ct < NN_FCT; ++ct)
                                    if (c0t + ct < nct)
                                        out.template operator()<0>((c0t + ct) * 16 + 4 * g, acc[ct], nullptr, pq[ct], z4);
                            }
                            // a width that is not a multiple of 32: the next layer's padding channels D .. D + 15 are 0
                            if (!last && (D & 31) && c0t == 0)
                                *(NCE_GLOBAL nn_bf16x4 *)(xout + (int64_t)t * m.xb_stride + D + 4 * g) = nn_bf16x4{};
                        }
                    }
                };
                if (m.cpad[l] == 64) layer.template operator()<2>();
                else layer.template operator()<1>();
                } else if (sizeof(Real) == 4 && MODE < 3 && (D & 7) == 0) {
                // projections in the scratch: eight channels per item (FP32), two 16-byte halves, so each hit's list
                // entry is read from memory once for both (with the projections in LDS four channels per item and
                // more hits in flight do better)
                using V4 = NNv4<Real>;
                const V4 z4 = V4{{(Real)0, (Real)0, (Real)0, (Real)0}};
                const int D8 = D >> 3;
                const int d8sh = (D8 & (D8 - 1)) == 0 ? __builtin_ctz((unsigned)D8) : -1;
                for (int k = tid; k < nb * D8; k += NN_BS) {
                    const int s = d8sh >= 0 ? k >> d8sh : k / D8, d = (k - s * D8) * 8;
                    const int64_t o = (int64_t)slot_site[s] * D + d;
                    const V4 pa = nn_ld4<Real>(cpre + o), pb = nn_ld4<Real>(cpre + o + 4);
                    V4 ga = z4, gb = z4;
                    if (m.has_gate) { ga = nn_ld4<Real>(cgpre + o); gb = nn_ld4<Real>(cgpre + o + 4); }
                    V4 dpa = z4, dpb = z4, dga = z4, dgb = z4;
                    if (s < prev) { dpa = nn_ld4<Real>(pr + (int64_t)s * M + d); dpb = nn_ld4<Real>(pr + (int64_t)s * M + d + 4); }
                    const int j0 = s == 0 ? 0 : cnt[s - 1], j1 = cnt[s];
                    auto at = [&](int h) -> int64_t { return (int64_t)(h >> 8) * M + D * (h & 0xff) + d; };
                    auto rd = [&](int64_t a) -> V4 { return nn_ld4<Real>(pr + a); };
                    int j = j0;
                    for (; j + 2 <= j1; j += 2) {
                        const int64_t a0 = at(list[j]), a1 = at(list[j + 1]);
                        const V4 x0 = rd(a0 + D), x1 = rd(a1 + D), x2 = rd(a0 + D + 4), x3 = rd(a1 + D + 4);
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            dpa.v[e] += x0.v[e] + x1.v[e];
                            dpb.v[e] += x2.v[e] + x3.v[e];
                        }
                        if (m.has_gate) {
                            const V4 y0 = rd(a0 + og), y1 = rd(a1 + og), y2 = rd(a0 + og + 4), y3 = rd(a1 + og + 4);
#pragma unroll
                            for (int e = 0; e < 4; ++e) {
                                dga.v[e] += y0.v[e] + y1.v[e];
                                dgb.v[e] += y2.v[e] + y3.v[e];
                            }
                        }
                    }
                    if (j < j1) {
                        const int64_t a = at(list[j]);
                        const V4 x0 = rd(a + D), x2 = rd(a + D + 4);
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            dpa.v[e] += x0.v[e];
                            dpb.v[e] += x2.v[e];
                        }
                        if (m.has_gate) {
                            const V4 y0 = rd(a + og), y2 = rd(a + og + 4);
#pragma unroll
                            for (int e = 0; e < 4; ++e) {
                                dga.v[e] += y0.v[e];
                                dgb.v[e] += y2.v[e];
                            }
                        }
                    }
                    V4 ha, hb;
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        dpa.v[e] += pa.v[e];                 // the new pre-activations
                        dpb.v[e] += pb.v[e];
                        dga.v[e] += ga.v[e];
                        dgb.v[e] += gb.v[e];
                        ha.v[e] = nn_act<Real>(m, dpa.v[e], dga.v[e]) - nn_act<Real>(m, pa.v[e], ga.v[e]);
                        hb.v[e] = nn_act<Real>(m, dpb.v[e], dgb.v[e]) - nn_act<Real>(m, pb.v[e], gb.v[e]);
                    }
                    // scratch rows by target: item k covers elements 8k .. 8k + 7
                    nn_st4<Real>(npre + 8 * (int64_t)k, dpa);
                    nn_st4<Real>(npre + 8 * (int64_t)k + 4, dpb);
                    if (m.has_gate) {
                        nn_st4<Real>(ngpre + 8 * (int64_t)k, dga);
                        nn_st4<Real>(ngpre + 8 * (int64_t)k + 4, dgb);
                    }
                    if (last) {
#pragma unroll
                        for (int e = 0; e < 4; ++e) part += (double)(ha.v[e] * hw[d + e]) + (double)(hb.v[e] * hw[d + 4 + e]);
                    } else {
                        nn_st4<Real>(dh + 8 * (int64_t)k, ha);
                        nn_st4<Real>(dh + 8 * (int64_t)k + 4, hb);
                    }
                }
                } else if (sizeof(Real) == 4 && (D & 3) == 0) {
                // four channels per item (FP32): every hit's projection block is read 16 bytes at a time
                using V4 = NNv4<Real>;
                constexpr int U2 = 2;                    // items (four channels each) per thread in flight together
                const int D4 = D >> 2;
                const int d4sh = (D4 & (D4 - 1)) == 0 ? __builtin_ctz((unsigned)D4) : -1;
                for (int k0 = tid; k0 < nb * D4; k0 += U2 * NN_BS) {
                    V4 p0[U2], g0[U2];
                    int sv[U2], dv[U2];
#pragma unroll
                    for (int u = 0; u < U2; ++u) {
                        const int k = k0 + u * NN_BS;
                        const int kk = k < nb * D4 ? k : 0;
                        sv[u] = d4sh >= 0 ? kk >> d4sh : kk / D4;
                        dv[u] = (kk - sv[u] * D4) * 4;
                        const int64_t o = (int64_t)slot_site[sv[u]] * D + dv[u];
                        p0[u] = nn_ld4<Real>(cpre + o);
                        if (m.has_gate) g0[u] = nn_ld4<Real>(cgpre + o);
                        else g0[u] = V4{{(Real)0, (Real)0, (Real)0, (Real)0}};
                    }
#pragma unroll
                    for (int u = 0; u < U2; ++u) {
                        const int k = k0 + u * NN_BS;
                        if (k >= nb * D4) break;
                        const int s = sv[u], d = dv[u];
                        V4 dp = s < prev ? nn_ld4<Real>(pr + (int64_t)s * M + d) : V4{{(Real)0, (Real)0, (Real)0, (Real)0}};
                        V4 dg = V4{{(Real)0, (Real)0, (Real)0, (Real)0}};
                        const int j0 = s == 0 ? 0 : cnt[s - 1], j1 = cnt[s];
                        auto at = [&](int h) -> int64_t { return (int64_t)(h >> 8) * M + D * (h & 0xff) + d; };
                        auto rd = [&](int64_t a) -> V4 { return nn_ld4<Real>(pr + a); };
                        int j = j0;
                        for (; j + 4 <= j1; j += 4) {
                            const int64_t a0 = at(list[j]), a1 = at(list[j + 1]), a2 = at(list[j + 2]),
                                          a3 = at(list[j + 3]);
                            const V4 x0 = rd(a0 + D), x1 = rd(a1 + D), x2 = rd(a2 + D), x3 = rd(a3 + D);
#pragma unroll
                            for (int e = 0; e < 4; ++e) dp.v[e] += (x0.v[e] + x1.v[e]) + (x2.v[e] + x3.v[e]);
                            if (m.has_gate) {
                                const V4 y0 = rd(a0 + og), y1 = rd(a1 + og), y2 = rd(a2 + og), y3 = rd(a3 + og);
#pragma unroll
                                for (int e = 0; e < 4; ++e) dg.v[e] += (y0.v[e] + y1.v[e]) + (y2.v[e] + y3.v[e]);
                            }
                        }
                        for (; j < j1; ++j) {
                            const int64_t a = at(list[j]);
                            const V4 x = rd(a + D);
#pragma unroll
                            for (int e = 0; e < 4; ++e) dp.v[e] += x.v[e];
                            if (m.has_gate) {
                                const V4 y = rd(a + og);
#pragma unroll
                                for (int e = 0; e < 4; ++e) dg.v[e] += y.v[e];
                            }
                        }
                        V4 dhv;
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            dp.v[e] += p0[u].v[e];               // the new pre-activations
                            dg.v[e] += g0[u].v[e];
                            dhv.v[e] = nn_act<Real>(m, dp.v[e], dg.v[e]) - nn_act<Real>(m, p0[u].v[e], g0[u].v[e]);
                        }
                        nn_st4<Real>(npre + 4 * (int64_t)k, dp);
                        if (m.has_gate) nn_st4<Real>(ngpre + 4 * (int64_t)k, dg);
                        if (last) {
#pragma unroll
                            for (int e = 0; e < 4; ++e) part += (double)(dhv.v[e] * hw[d + e]);
                        } else {
                            nn_st4<Real>(dh + 4 * (int64_t)k, dhv);
                        }
                    }
                }
                } else if ((D & 1) == 0) {
                // two channels per item (FP64, or FP32 with an even width): each hit's projection block is read
                // two values per load
                using V2 = NNv2<Real>;
                constexpr int U2 = 2;                    // items (two channels each) per thread in flight together
                const int D2 = D >> 1;
                const int d2sh = (D2 & (D2 - 1)) == 0 ? __builtin_ctz((unsigned)D2) : -1;
                for (int k0 = tid; k0 < nb * D2; k0 += U2 * NN_BS) {
                    V2 p0[U2], g0[U2];
                    int sv[U2], dv[U2];
#pragma unroll
                    for (int u = 0; u < U2; ++u) {
                        const int k = k0 + u * NN_BS;
                        const int kk = k < nb * D2 ? k : 0;
                        sv[u] = d2sh >= 0 ? kk >> d2sh : kk / D2;
                        dv[u] = (kk - sv[u] * D2) * 2;
                        const int64_t o = (int64_t)slot_site[sv[u]] * D + dv[u];
                        p0[u] = nn_ld2<Real>(cpre + o);
                        if (m.has_gate) g0[u] = nn_ld2<Real>(cgpre + o);
                        else g0[u] = V2{{(Real)0, (Real)0}};
                    }
#pragma unroll
                    for (int u = 0; u < U2; ++u) {
                        const int k = k0 + u * NN_BS;
                        if (k >= nb * D2) break;
                        const int s = sv[u], d = dv[u];
                        V2 dp = s < prev ? nn_ld2<Real>(pr + (int64_t)s * M + d) : V2{{(Real)0, (Real)0}};
                        V2 dg = V2{{(Real)0, (Real)0}};
                        const int j0 = s == 0 ? 0 : cnt[s - 1], j1 = cnt[s];
                        auto at = [&](int h) -> int64_t { return (int64_t)(h >> 8) * M + D * (h & 0xff) + d; };
                        auto rd = [&](int64_t a) -> V2 { return nn_ld2<Real>(pr + a); };
                        int j = j0;
                        for (; j + 4 <= j1; j += 4) {
                            const int64_t a0 = at(list[j]), a1 = at(list[j + 1]), a2 = at(list[j + 2]),
                                          a3 = at(list[j + 3]);
                            const V2 x0 = rd(a0 + D), x1 = rd(a1 + D), x2 = rd(a2 + D), x3 = rd(a3 + D);
#pragma unroll
                            for (int e = 0; e < 2; ++e) dp.v[e] += (x0.v[e] + x1.v[e]) + (x2.v[e] + x3.v[e]);
                            if (m.has_gate) {
                                const V2 y0 = rd(a0 + og), y1 = rd(a1 + og), y2 = rd(a2 + og), y3 = rd(a3 + og);
#pragma unroll
                                for (int e = 0; e < 2; ++e) dg.v[e] += (y0.v[e] + y1.v[e]) + (y2.v[e] + y3.v[e]);
                            }
                        }
                        for (; j < j1; ++j) {
                            const int64_t a = at(list[j]);
                            const V2 x = rd(a + D);
#pragma unroll
                            for (int e = 0; e < 2; ++e) dp.v[e] += x.v[e];
                            if (m.has_gate) {
                                const V2 y = rd(a + og);
#pragma unroll
                                for (int e = 0; e < 2; ++e) dg.v[e] += y.v[e];
                            }
                        }
                        V2 dhv;
#pragma unroll
                        for (int e = 0; e < 2; ++e) {
                            dp.v[e] += p0[u].v[e];               // the new pre-activations
                            dg.v[e] += g0[u].v[e];
                            dhv.v[e] = nn_act<Real>(m, dp.v[e], dg.v[e]) - nn_act<Real>(m, p0[u].v[e], g0[u].v[e]);
                        }
                        nn_st2<Real>(npre + 2 * (int64_t)k, dp);
                        if (m.has_gate) nn_st2<Real>(ngpre + 2 * (int64_t)k, dg);
                        if (last) {
#pragma unroll
                            for (int e = 0; e < 2; ++e) part += (double)(dhv.v[e] * hw[d + e]);
                        } else {
                            nn_st2<Real>(dh + 2 * (int64_t)k, dhv);
                        }
                    }
                }
                } else
                for (int k0 = tid; k0 < nb * D; k0 += U * NN_BS) {
                    Real p0[U], g0[U];
                    int sv[U], dv[U];
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const int k = k0 + u * NN_BS;
                        const int kk = k < nb * D ? k : 0;
                        sv[u] = kk / D; dv[u] = kk - sv[u] * D;
                        const int64_t o = (int64_t)slot_site[sv[u]] * D + dv[u];
                        p0[u] = cpre[o];
                        g0[u] = m.has_gate ? cgpre[o] : (Real)0;
                    }
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const int k = k0 + u * NN_BS;
                        if (k >= nb * D) break;
                        const int s = sv[u], d = dv[u];
                        Real dp = s < prev ? pr[(int64_t)s * M + d] : (Real)0;
                        Real dg = (Real)0;
                        const int j0 = s == 0 ? 0 : cnt[s - 1], j1 = cnt[s];
                        // hits four at a time: independent loads in flight together
                        auto at = [&](int h) -> int64_t { return (int64_t)(h >> 8) * M + D * (h & 0xff) + d; };
                        int j = j0;
                        for (; j + 4 <= j1; j += 4) {
                            const int64_t a0 = at(list[j]), a1 = at(list[j + 1]), a2 = at(list[j + 2]),
                                          a3 = at(list[j + 3]);
                            dp += (pr[a0 + D] + pr[a1 + D]) + (pr[a2 + D] + pr[a3 + D]);
                            if (m.has_gate) dg += (pr[a0 + og] + pr[a1 + og]) + (pr[a2 + og] + pr[a3 + og]);
                        }
                        for (; j < j1; ++j) {
                            const int64_t a = at(list[j]);
                            dp += pr[a + D];
                            if (m.has_gate) dg += pr[a + og];
                        }
                        const Real np = p0[u] + dp, ng = g0[u] + dg;   // the new pre-activations
                        const Real dhv = nn_act<Real>(m, np, ng) - nn_act<Real>(m, p0[u], g0[u]);
                        npre[k] = np;
                        if (m.has_gate) ngpre[k] = ng;
                        if (last) part += (double)(dhv * hw[d]);
                        else dh[k] = dhv;
                    }
                }
                if constexpr (MFMA) {
                    if (l + 1 < L) gbar();               // this layer's output rows are in place
                } else {
                    __syncthreads();
                }
            }
            w_staged = L == 1;                           // one layer: its matrix stays in LDS for the launch
            double diff;
            if constexpr (SPLIT) {
                // every part's partial sum, then the same sum in every part (fixed order)
                const double mine = nn_block_sum<double>(part, red);
                if (tid == 0) sp.part[(int64_t)item * S + prt] = mine;
                gbar();
                diff = 0.0;
                for (int q = 0; q < S; ++q) diff += sp.part[(int64_t)item * S + q];
            } else {
                diff = nn_block_sum<double>(part, red);
            }
            const double dE = (m.readout == 0 ? diff / (double)m.nA : diff);
            const double dE_eV = dE * m.de_scale;
            // evaluation only (nce_nn_delta): report the change, leave every state buffer untouched
            const bool accept = eval_dE == nullptr && ((dE_eV <= 0.0) || (u < exp(-beta * dE_eV)));
            if (eval_dE && tid == 0) eval_dE[blockIdx.x] = dE;
            if (accept) {
                // (the parts of a split replica update alternate rows)
                // (NN_CU items per thread per round: every cache load of a round is issued before its first add, so
                // the rounds, not the items, pay the memory latency)
                if (m.cache_h0) {
                    const int nk = n_own(counts[0]) * C0;
                    for (int k0 = tid; k0 < nk; k0 += NN_CU * NN_BS) {
                        int64_t o[NN_CU];
                        Real cv[NN_CU], dv[NN_CU];
#pragma unroll
                        for (int u = 0; u < NN_CU; ++u) {
                            const int k = min(k0 + u * NN_BS, nk - 1);
                            const int so = k / C0, c = k - so * C0, s = prt + so * S;
                            o[u] = (int64_t)slot_site[s] * C0 + c;
                            dv[u] = dh0[(int64_t)s * C0 + c];
                            cv[u] = cache[o[u]];
                        }
#pragma unroll
                        for (int u = 0; u < NN_CU; ++u)
                            if (k0 + u * NN_BS < nk) cache[o[u]] = cv[u] + dv[u];
                    }
                }
                for (int l = 0; l < L; ++l) {
                    const int D = m.dd[l];
                    // the new pre-activations of the step's targets replace the cached rows (the same FP32 or FP64
                    // sums the layers formed)
                    const NCE_GLOBAL Real *npre = scr + m.soff_npre[l];
                    const NCE_GLOBAL Real *ngpre = scr + m.soff_ngpre[l];
                    if (sizeof(Real) == 4 && (D & 3) == 0) {
                        // four channels per thread: 16-byte rows of the scratch and the cache
                        using V4 = NNv4<Real>;
                        const int D4 = D >> 2;
                        const int d4sh = (D4 & (D4 - 1)) == 0 ? __builtin_ctz((unsigned)D4) : -1;
                        for (int g = 0; g < (m.has_gate ? 2 : 1); ++g) {
                            NCE_GLOBAL Real *cbase = cache + (g == 0 ? m.off_pre[l] : m.off_gpre[l]);
                            const NCE_GLOBAL Real *nbase = g == 0 ? npre : ngpre;
                            const int nk = n_own(counts[l + 1]) * D4;
                            for (int k0 = tid; k0 < nk; k0 += NN_CU * NN_BS) {
                                NCE_GLOBAL Real *cp[NN_CU];
                                V4 nv[NN_CU];
#pragma unroll
                                for (int u = 0; u < NN_CU; ++u) {
                                    const int k = min(k0 + u * NN_BS, nk - 1);
                                    const int so = d4sh >= 0 ? k >> d4sh : k / D4, d = (k - so * D4) * 4;
                                    const int s = prt + so * S;
                                    cp[u] = cbase + (int64_t)slot_site[s] * D + d;
                                    nv[u] = nn_ld4<Real>(nbase + (int64_t)s * D + d);
                                }
#pragma unroll
                                for (int u = 0; u < NN_CU; ++u)
                                    if (k0 + u * NN_BS < nk) nn_st4<Real>(cp[u], nv[u]);
                            }
                        }
                    } else {
                        const int dsh = (D & (D - 1)) == 0 ? __builtin_ctz((unsigned)D) : -1;
                        const int nk = n_own(counts[l + 1]) * D;
                        for (int g = 0; g < (m.has_gate ? 2 : 1); ++g) {
                            NCE_GLOBAL Real *cbase = cache + (g == 0 ? m.off_pre[l] : m.off_gpre[l]);
                            const NCE_GLOBAL Real *nbase = g == 0 ? npre : ngpre;
                            for (int k0 = tid; k0 < nk; k0 += NN_CU * NN_BS) {
                                int64_t o[NN_CU];
                                Real nv[NN_CU];
#pragma unroll
                                for (int u = 0; u < NN_CU; ++u) {
                                    const int k = min(k0 + u * NN_BS, nk - 1);
                                    const int so = dsh >= 0 ? k >> dsh : k / D, d = k - so * D, s = prt + so * S;
                                    o[u] = (int64_t)slot_site[s] * D + d;
                                    nv[u] = nbase[(int64_t)s * D + d];
                                }
#pragma unroll
                                for (int u = 0; u < NN_CU; ++u)
                                    if (k0 + u * NN_BS < nk) cbase[o[u]] = nv[u];
                            }
                        }
                    }
                }
                if (tid == 0) {
                    occ[ia] = (int8_t)sj;
                    occ[ib] = (int8_t)si;
                }
                // later proposals of the batch see the swap: occ[ia] is now sj, occ[ib] is si
                if (tid > la_slot && tid < NN_LA) {
                    NNProp &q = s_la[tid];
                    q.si = q.ia == ia ? sj : q.ia == ib ? si : q.si;
                    q.sj = q.ib == ia ? sj : q.ib == ib ? si : q.sj;
                }
                E += dE;
                ++nacc;
            }
            // clear the table for the next step (a step on the fixed tables leaves it empty)
            if (!fixed) {
                for (int i = tid; i < hcap; i += NN_BS) keys[i] = NN_EMPTY;
                if (tid == 0) misc[0] = 0;
            }
        }
        if (job.traj && eval_dE == nullptr && ((t + 1) % job.sample_interval) == 0 && tid == 0 && prt == 0)
            job.traj[((t + 1) / job.sample_interval - 1) * job.n_replicas + r] = E;
        __syncthreads();
    }
    if (tid == 0 && eval_dE == nullptr && prt == 0) {
        // a split replica whose parts lost each other returns NaN (its steps ran without waiting)
        if (SPLIT && misc[1]) E = __builtin_nan("");
        job.energy[r] = E;
        job.n_accept[r] += nacc;
        job.step[r] = (int64_t)(step0 + (uint64_t)n_steps);
    }
}

// barrier counters of a split launch start at zero (a fill kernel: the launch may be captured, and memset nodes in
// captured graphs are avoided)
__global__ void nn_zero_u32(uint32_t *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// ── host side ──────────────────────────────────────────────────────────────────────────────────────────────
static int nn_sync_models(nce_ctx *c) {
    if ((int)c->nn_models.size() > c->cap_nn_models) {
        const int cap = std::max(8, 2 * (int)c->nn_models.size());
        void *p = nullptr;
        HIP_TRY(hipMalloc(&p, cap * sizeof(NNModelDev)));
        if (c->d_nn_models) HIP_TRY(hipFree(c->d_nn_models));
        c->d_nn_models = p;
        c->cap_nn_models = cap;
    }
    HIP_TRY(hipMemcpy(c->d_nn_models, c->nn_models.data(), c->nn_models.size() * sizeof(NNModelDev),
                      hipMemcpyHostToDevice));
    return NCE_OK;
}

// LDS placement of the chunk kernel's per-step arrays: the most that fits beside the hash table, in the order
// layer-0 change, partner hits, projections
// (the BF16 layers keep only the layer-0 change there: at most mode 1)
template <typename Real>
static int nn_lds_mode(const NNModelDev &m, bool fused = false) {
    int mode = fused ? 1 : 3;
    while (mode > 0 && nn_lds_for<Real>(m, mode, fused).total > 160 * 1024) --mode;
    return mode;
}

// float -> bf16 bits, round to nearest even (the device conversion v_cvt_pk_bf16_f32 rounds the same way)
static uint16_t nn_bf16_bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// The stacked matrix of BF16 layer l in its LDS layout: one row per output column j (gated models alternate 16-column
// tiles of pre-activations and gates), Kf + 8 elements per row, k = segment * cpad + c for input channel c of the
// segments self, then shell p's sum as a high (1 + 2p) and a low (2 + 2p) part.  Wc: the layer's [cin][mcat] matrix
// in float.
static void nn_stack_bf16(const NNModelDev &m, int l, const float *Wc, const int32_t *c0_ch, std::vector<uint16_t> &img) {
    const int D = m.dd[l], M = m.mcat[l], P = m.P, Cp = m.cpad[l], Kf = m.kf[l], Ks = Kf + 8;
    const int Cin = l == 0 ? m.c0_eff : m.cin[l];
    const int Nout = D * (m.has_gate ? 2 : 1);
    for (int j = 0; j < Nout; ++j)
        for (int kk = 0; kk < Ks; ++kk) {
            float v = 0.0f;
            if (kk < Kf) {
                const int seg = kk / Cp, c = kk - seg * Cp;
                const int T = j >> 4;
                const bool isg = m.has_gate && (T & 1);
                const int dch = m.has_gate ? (T >> 1) * 16 + (j & 15) : j;
                if (c < Cin) {
                    const int64_t row = (int64_t)(l == 0 ? c0_ch[c] : c) * M;
                    if (seg == 0) {
                        if (!isg) v = Wc[row + dch];
                    } else {
                        const int p = (seg - 1) >> 1;
                        v = Wc[row + (isg ? D * (1 + P + p) : D * (1 + p)) + dch];
                    }
                }
            }
            img.push_back(nn_bf16_bits(v));
        }
}

extern "C" int nce_nn_model_create(nce_ctx *c, const nce_nn_model_desc *d, int32_t *model_id) {
    if (!c || !d || !model_id) return set_err(NCE_ERR_ARG, "NULL argument");
    HIP_TRY(hipSetDevice(c->device));
    if (d->n_layers < 0 || d->n_layers > NN_MAXL - 1) return set_err(NCE_ERR_ARG, "0..7 hidden layers supported");
    if (d->n_window < 1 || d->n_window >= 0xFFFF || d->n_entries < 1 || d->c0 < 1 || d->n_shells < 0 ||
        d->n_shells > 255)
        return set_err(NCE_ERR_ARG, "model sizes out of range");
    for (int i = 0; i < 3; ++i)
        if (d->d[i] < 1) return set_err(NCE_ERR_ARG, "cell diagonal must be positive");
    if ((int64_t)d->d[0] * d->d[1] * d->d[2] != d->n_cells || d->n_cells >= ((int64_t)1 << 31))
        return set_err(NCE_ERR_ARG, "n_cells inconsistent or too large");
    NNModelDev m{};
    m.d0 = d->d[0]; m.d1 = d->d[1]; m.d2 = d->d[2];
    m.n_cells = d->n_cells; m.n_sites = d->n_cells; m.nA = (int32_t)d->n_cells;
    magic_u32((uint32_t)(m.d1 * m.d2), &m.div12_m, &m.div12_s);
    magic_u32((uint32_t)m.d2, &m.div2_m, &m.div2_s);
    m.W0 = d->n_window; m.J0 = d->n_entries; m.P = d->n_shells; m.C0 = d->c0; m.L = d->n_layers;
    m.act = d->activation; m.readout = d->readout; m.has_gate = d->w_gate != nullptr;
    if (d->cache_mode != 0 && d->cache_mode != 1) return set_err(NCE_ERR_ARG, "cache_mode must be 0 (h0) or 1 (proj)");
    m.cache_h0 = d->cache_mode == 0 || d->n_layers == 0;
    m.head_b = d->head_b; m.de_scale = d->de_scale;
    int n_part = 0;
    for (int p = 0; p < m.P; ++p) {
        if (d->shell_count[p] < 0) return set_err(NCE_ERR_ARG, "negative shell count");
        n_part += d->shell_count[p];
    }
    m.n_part = n_part;
    const int64_t N = m.n_sites;
    // weights: per layer self [cin][d], shell [P][cin][d], gate [P][cin][d], b [d], gb [d], then the projection
    // matrix [cin][mcat] = self | shell_p / n_p | gate_p / n_p
    std::vector<double> wts;
    int64_t so = 0, sho = 0, go = 0, bo = 0, gbo = 0;
    int cin_prev = m.C0;
    std::vector<int32_t> c0_ch;
    for (int l = 0; l < m.L; ++l) {
        const int C = d->layer_cin[l], D = d->layer_d[l];
        if (C != cin_prev || D < 1) return set_err(NCE_ERR_ARG, "layer widths do not chain");
        m.cin[l] = C; m.dd[l] = D;
        const int M = D * (1 + m.P + (m.has_gate ? m.P : 0));
        m.mcat[l] = M;
        const double *ws = d->w_self + so, *wsh = d->w_shell + sho, *wg = m.has_gate ? d->w_gate + go : nullptr;
        m.off_self[l] = (int64_t)wts.size();
        wts.insert(wts.end(), ws, ws + (int64_t)C * D);
        m.off_shell[l] = (int64_t)wts.size();
        wts.insert(wts.end(), wsh, wsh + (int64_t)m.P * C * D);
        m.off_gate[l] = (int64_t)wts.size();
        if (m.has_gate) wts.insert(wts.end(), wg, wg + (int64_t)m.P * C * D);
        m.off_b[l] = (int64_t)wts.size();
        wts.insert(wts.end(), d->b + bo, d->b + bo + D);
        m.off_gb[l] = (int64_t)wts.size();
        if (m.has_gate) wts.insert(wts.end(), d->gb + gbo, d->gb + gbo + D);
        m.off_wcat[l] = (int64_t)wts.size();
        std::vector<double> wc((size_t)C * M, 0.0);
        for (int cc = 0; cc < C; ++cc) {
            for (int dd = 0; dd < D; ++dd) wc[(size_t)cc * M + dd] = ws[(int64_t)cc * D + dd];
            for (int p = 0; p < m.P; ++p) {
                const double inv = d->shell_count[p] > 0 ? 1.0 / d->shell_count[p] : 0.0;
                for (int dd = 0; dd < D; ++dd) {
                    wc[(size_t)cc * M + D * (1 + p) + dd] = wsh[((int64_t)p * C + cc) * D + dd] * inv;
                    if (m.has_gate) wc[(size_t)cc * M + D * (1 + m.P + p) + dd] = wg[((int64_t)p * C + cc) * D + dd] * inv;
                }
            }
        }
        if (l == 0)
            for (int cc = 0; cc < C; ++cc) {
                bool nz = false;
                for (int j = 0; j < M; ++j) nz |= wc[(size_t)cc * M + j] != 0.0;
                if (nz) c0_ch.push_back(cc);
            }
        wts.insert(wts.end(), wc.begin(), wc.end());
        so += (int64_t)C * D; sho += (int64_t)m.P * C * D; bo += D;
        if (m.has_gate) { go += (int64_t)m.P * C * D; gbo += D; }
        cin_prev = D;
    }
    if (wts.empty()) wts.push_back(0.0);
    if (c0_ch.empty()) c0_ch.push_back(0);
    m.c0_eff = m.L > 0 ? (int32_t)c0_ch.size() : 0;
    const int DL = m.L > 0 ? m.dd[m.L - 1] : 0;
    // slots
    // the entries interleaved over the channels (the i-th entry of every channel, then the (i+1)-th, ...): every
    // entry adds to the row of the flip's own site, and consecutive entries of one channel would make the lanes of
    // one atomic add meet at one address.  A channel's entries keep their order, so the full recompute, which adds
    // them in table order, sums them as before.
    for (int e = 0; e < m.J0; ++e)
        if (d->entry_col[e] < 1 || d->entry_col[e] >= m.C0) return set_err(NCE_ERR_ARG, "entry channel out of range");
    std::vector<int> eord(m.J0), erank(m.J0), eseen(m.C0, 0);
    for (int e = 0; e < m.J0; ++e) {
        eord[e] = e;
        erank[e] = eseen[d->entry_col[e]]++;
    }
    std::stable_sort(eord.begin(), eord.end(), [&](int a, int b) {
        return erank[a] != erank[b] ? erank[a] < erank[b] : d->entry_col[a] < d->entry_col[b];
    });
    std::vector<uint2> slots(m.J0);
    std::vector<int32_t> ent_col_s(std::max(m.J0, 1), 0);
    std::vector<double> ent_coef_s(std::max(m.J0, 1), 0.0);
    for (int i = 0; i < m.J0; ++i) {
        const int e = eord[i];
        uint32_t q[4];
        for (int k = 0; k < 4; ++k) {
            q[k] = d->entry_slot[(size_t)e * 4 + k];
            if (q[k] != 0xFFFFu && (int)q[k] >= m.W0) return set_err(NCE_ERR_ARG, "entry slot outside the window");
        }
        slots[i] = make_uint2(q[0] | (q[1] << 16), q[2] | (q[3] << 16));
        ent_col_s[i] = d->entry_col[e];
        ent_coef_s[i] = d->entry_coef[e];
    }
    auto chk4 = [&](const int32_t *w, int n) {
        for (int i = 0; i < n; ++i)
            if (w[4 * i] != 0 || w[4 * i + 1] < 0 || w[4 * i + 1] >= m.d0 || w[4 * i + 2] < 0 || w[4 * i + 2] >= m.d1 ||
                w[4 * i + 3] < 0 || w[4 * i + 3] >= m.d2)
                return false;
        return true;
    };
    if (!chk4(d->window, m.W0) || !chk4(d->partners, n_part)) return set_err(NCE_ERR_ARG, "offsets not reduced");
    std::vector<int32_t> part_shell(std::max(n_part, 1)), part_col(std::max(n_part, 1));
    int max_sc = 1;
    for (int p = 0; p < m.P; ++p) max_sc = std::max(max_sc, (int)d->shell_count[p]);
    m.nbr_ps = (max_sc + 3) & ~3;
    m.nbr_w = std::max(1, m.P) * m.nbr_ps;
    for (int p = 0, k = 0; p < m.P; ++p)
        for (int q = 0; q < d->shell_count[p]; ++q) {
            part_col[k] = p * m.nbr_ps + q;
            part_shell[k++] = p;
        }
    // balls: the sites of one flip's ball_l on this cell (window, then dilated by the partner offsets l times, as
    // the kernel builds it); the union of two flips holds at most twice as many, and never more than the cell
    {
        auto site_of = [&](int64_t centre, const int32_t *o) {
            const int64_t k2c = centre % m.d2, k1c = (centre / m.d2) % m.d1, k0c = centre / ((int64_t)m.d1 * m.d2);
            const int64_t k0 = (k0c + o[1]) % m.d0, k1 = (k1c + o[2]) % m.d1, k2 = (k2c + o[3]) % m.d2;
            return (k0 * m.d1 + k1) * m.d2 + k2;
        };
        std::unordered_set<int64_t> ball;
        std::vector<int64_t> cur;
        int64_t n1[NN_MAXL + 1];
        for (int k = 0; k < m.W0; ++k)
            if (ball.insert(site_of(0, d->window + 4 * k)).second) cur.push_back(site_of(0, d->window + 4 * k));
        for (int l = 0; l <= m.L; ++l) {
            if (l > 0) {
                const std::vector<int64_t> prev(cur);
                for (int64_t x : prev)
                    for (int k = 0; k < n_part; ++k) {
                        const int64_t y = site_of(x, d->partners + 4 * k);
                        if (ball.insert(y).second) cur.push_back(y);
                    }
            }
            n1[l] = (int64_t)cur.size();
            m.ball_cap[l] = (int32_t)std::min<int64_t>(2 * (int64_t)cur.size(), N);
            if ((int64_t)m.ball_cap[l] >= ((int64_t)1 << 23))
                return set_err(NCE_ERR_UNSUPPORTED, "receptive field too large");
        }
        // Fixed-ball tables.  cur lists one flip's ball_L around site 0, each ball_l a prefix of n1[l] sites; by
        // translation the ball of any centre c is c + cur.  The union of two disjoint balls numbers its slots by
        // layer, flip 1 before flip 2 within a layer, so ball_l is the prefix 2 n1[l] of the union as the kernel
        // requires.  The tables are skipped where two balls cannot be apart on this cell or are too large to pair.
        const char *fixed_env = std::getenv("NCE_NN_FIXED_BALLS");
        const int64_t nL = n1[m.L];
        if (m.L > 0 && n_part > 0 && 2 * nL <= N && nL <= 8192 && !(fixed_env && fixed_env[0] == '0')) {
            std::unordered_map<int64_t, int32_t> idx;
            for (int64_t j = 0; j < nL; ++j) idx.emplace(cur[j], (int32_t)j);
            auto uslot = [&](int f, int64_t j) -> int32_t {
                int l = 0;
                while (j >= n1[l]) ++l;
                const int64_t lo = l > 0 ? n1[l - 1] : 0;
                return (int32_t)(2 * lo + f * (n1[l] - lo) + (j - lo));
            };
            auto coords = [&](int64_t s, int64_t &k0, int64_t &k1, int64_t &k2) {
                k2 = s % m.d2; k1 = (s / m.d2) % m.d1; k0 = s / ((int64_t)m.d1 * m.d2);
            };
            std::vector<int4> slot(2 * nL);
            std::vector<int32_t> src_f(2 * nL), src_j(2 * nL);
            for (int f = 0; f < 2; ++f)
                for (int64_t j = 0; j < nL; ++j) {
                    int64_t k0, k1, k2;
                    coords(cur[j], k0, k1, k2);
                    const int32_t u = uslot(f, j);
                    slot[u] = make_int4(f, (int)k0, (int)k1, (int)k2);
                    src_f[u] = f;
                    src_j[u] = (int32_t)j;
                }
            std::vector<int32_t> win_slot(2 * (size_t)m.W0);
            for (int f = 0; f < 2; ++f)
                for (int k = 0; k < m.W0; ++k) win_slot[(size_t)f * m.W0 + k] = uslot(f, idx.at(site_of(0, d->window + 4 * k)));
            std::vector<int32_t> list, ends;
            // the zero row of the BF16 layers' input buffers (xb_zrow): one past the largest ball_0 .. ball_{L-1}
            int64_t nbr_zrow = 1;
            for (int l = 0; l < m.L; ++l) nbr_zrow = std::max<int64_t>(nbr_zrow, m.ball_cap[l]);
            std::vector<uint16_t> nbr;
            const bool nbr_ok = nbr_zrow < 0xFFFF;
            for (int l = 0; l < m.L && nbr_ok; ++l) {
                // per target of ball_l: the source slot of each partner column, the zero row where none
                const int64_t nsrc = 2 * n1[l], ntgt = 2 * n1[l + 1];
                m.pre_nbr_off[l] = (int64_t)nbr.size();
                nbr.resize(nbr.size() + (size_t)(ntgt * m.nbr_w), (uint16_t)nbr_zrow);
                uint16_t *np = nbr.data() + m.pre_nbr_off[l];
                for (int64_t q = 0; q < nsrc; ++q)
                    for (int k = 0; k < n_part; ++k) {
                        const int64_t y = site_of(cur[src_j[q]], d->partners + 4 * k);
                        np[(size_t)(uslot(src_f[q], idx.at(y)) * m.nbr_w + part_col[k])] = (uint16_t)q;
                    }
            }
            if (nbr_ok) UP(nbr.data(), nbr.size(), &m.pre_nbr);
            for (int l = 0; l < m.L; ++l) {
                // hits in the kernel's order (source q, partner k), counted by target, then placed stably
                const int64_t nsrc = 2 * n1[l], ntgt = 2 * n1[l + 1];
                std::vector<int32_t> tgt((size_t)(nsrc * n_part)), start((size_t)ntgt + 1, 0);
                for (int64_t q = 0; q < nsrc; ++q)
                    for (int k = 0; k < n_part; ++k) {
                        const int64_t y = site_of(cur[src_j[q]], d->partners + 4 * k);
                        const int32_t t = uslot(src_f[q], idx.at(y));
                        tgt[(size_t)(q * n_part + k)] = t;
                        ++start[(size_t)t + 1];
                    }
                for (int64_t t = 0; t < ntgt; ++t) start[(size_t)t + 1] += start[(size_t)t];
                m.pre_end_off[l] = (int64_t)ends.size();
                ends.insert(ends.end(), start.begin() + 1, start.end());
                m.pre_list_off[l] = (int64_t)list.size();
                list.resize(list.size() + (size_t)(nsrc * n_part));
                int32_t *lp = list.data() + m.pre_list_off[l];
                for (int64_t q = 0; q < nsrc; ++q)
                    for (int k = 0; k < n_part; ++k)
                        lp[start[(size_t)tgt[(size_t)(q * n_part + k)]]++] = (int32_t)(q << 8) | part_shell[k];
            }
            // ball_L - ball_L on the cell
            std::vector<uint32_t> near((size_t)((N + 31) / 32), 0u);
            std::vector<int64_t> c0(nL), c1(nL), c2(nL);
            for (int64_t j = 0; j < nL; ++j) coords(cur[j], c0[j], c1[j], c2[j]);
            for (int64_t a = 0; a < nL; ++a)
                for (int64_t b = 0; b < nL; ++b) {
                    const int64_t v0 = (c0[a] - c0[b] + m.d0) % m.d0, v1 = (c1[a] - c1[b] + m.d1) % m.d1,
                                  v2 = (c2[a] - c2[b] + m.d2) % m.d2;
                    const int64_t v = (v0 * m.d1 + v1) * m.d2 + v2;
                    near[(size_t)(v >> 5)] |= 1u << (v & 31);
                }
            for (int l = 0; l <= m.L; ++l) m.pre_n[l] = (int32_t)n1[l];
            UP(near.data(), near.size(), &m.pre_near);
            UP(slot.data(), slot.size(), &m.pre_slot);
            UP(win_slot.data(), win_slot.size(), &m.pre_win);
            UP(list.data(), list.size(), &m.pre_list);
            UP(ends.data(), ends.size(), &m.pre_end);
            m.pre_ok = 1;
        }
    }
    int hc = 64;
    while (hc < m.ball_cap[m.L] + m.ball_cap[m.L] / 2) hc <<= 1;
    m.hash_cap = hc;
    m.hash_cap_f = hc;                           // grown below once the fused layout's placement is known
    // per-replica cache layout; every region starts at a multiple of 4 elements (16-byte rows of four channels)
    auto pad4 = [](int64_t v) { return (v + 3) & ~(int64_t)3; };
    int64_t o = m.cache_h0 ? pad4(N * m.C0) : 0;
    for (int l = 0; l < m.L; ++l) {
        m.off_pre[l] = o; o = pad4(o + N * m.dd[l]);
        m.off_gpre[l] = o; if (m.has_gate) o = pad4(o + N * m.dd[l]);
    }
    m.cache_elems = o;
    // per-replica scratch layout (elements of the precision's type; int arrays use one element per int)
    int64_t s2 = 0, pr_max = 1, hit_max = 1, dh_max = 1;
    for (int l = 0; l < m.L; ++l) {
        pr_max = std::max(pr_max, (int64_t)m.ball_cap[l] * m.mcat[l]);
        hit_max = std::max(hit_max, (int64_t)m.ball_cap[l] * std::max(n_part, 1));
        dh_max = std::max(dh_max, (int64_t)m.ball_cap[l + 1] * m.dd[l]);
    }
    // BF16 layers: input widths padded to 32, stacked matrices, and two input-row buffers of bf16 rows in the
    // projection block (rows 16-byte aligned, the second buffer 128-byte aligned)
    m.fused_ok = m.L > 0;
    m.xb_stride = 8;
    m.wf_bytes = 16;
    int64_t xb_rows = 1;
    for (int l = 0; l < m.L; ++l) {
        const int cin = l == 0 ? m.c0_eff : m.cin[l];
        m.cpad[l] = (cin + 31) & ~31;
        m.kf[l] = m.cpad[l] * (1 + 2 * m.P);
        m.xb_stride = std::max(m.xb_stride, m.cpad[l]);
        const int nout = m.dd[l] * (m.has_gate ? 2 : 1);
        m.wf_bytes = std::max<int64_t>(m.wf_bytes, (int64_t)nout * (m.kf[l] + 8) * 2);
        xb_rows = std::max<int64_t>(xb_rows, m.ball_cap[l]);
        if (m.dd[l] % 16 != 0 || nout > 16 * NN_FMAXT || m.cpad[l] > 64) m.fused_ok = 0;
    }
    if (xb_rows >= 0xFFFF) m.fused_ok = 0;       // slots and the zero row must fit the neighbour tables' uint16
    m.xb_zrow = (int32_t)xb_rows;                // one row past the largest ball, zeroed by every launch
    m.xb_off1 = ((xb_rows + 1) * m.xb_stride + 63) & ~(int64_t)63;
    if (m.fused_ok && m.pre_ok && !m.pre_nbr) m.fused_ok = 0;
    // two bf16 buffers of xb_off1 elements each take xb_off1 float elements
    const int64_t pr_region = m.fused_ok ? std::max<int64_t>(pr_max, m.xb_off1 + 64) : pr_max;
    auto pad32 = [](int64_t v) { return (v + 31) & ~(int64_t)31; };
    m.soff_dh0 = s2; s2 = pad32(s2 + (int64_t)m.ball_cap[0] * m.C0);
    m.soff_pr = s2; s2 = pad4(s2 + pr_region);
    m.soff_tgt = s2; s2 = pad4(s2 + hit_max);
    // BF16 layers: the neighbour table of a hash step (ball_cap[L] rows of nbr_w uint16, every layer's) and the slots
    // a split replica's builder publishes (counts, window slots, sites)
    m.soff_nbr = s2;
    if (m.fused_ok) s2 = pad4(s2 + ((int64_t)m.ball_cap[m.L] * m.nbr_w + 1) / 2);
    m.soff_pub = s2; if (m.fused_ok) s2 = pad4(s2 + m.L + 1 + 2 * (int64_t)m.W0 + m.ball_cap[m.L]);
    m.soff_list = s2; s2 = pad4(s2 + hit_max);
    for (int l = 0; l < m.L; ++l) {
        m.soff_npre[l] = s2; s2 = pad4(s2 + (int64_t)m.ball_cap[l + 1] * m.dd[l]);
        m.soff_ngpre[l] = s2; if (m.has_gate) s2 = pad4(s2 + (int64_t)m.ball_cap[l + 1] * m.dd[l]);
    }
    m.soff_dh[0] = s2; s2 = pad4(s2 + dh_max);
    m.soff_dh[1] = s2; s2 = pad4(s2 + dh_max);
    m.scratch_elems = pad32(s2);                 // every replica's scratch starts 128-byte aligned
    std::vector<float> mu_f(m.C0), sd_f(m.C0), lin_f(m.C0), w_f(wts.size()), hw_f(std::max(DL, 1));
    for (int i = 0; i < m.C0; ++i) { mu_f[i] = (float)d->mu[i]; sd_f[i] = (float)d->sd[i]; lin_f[i] = (float)d->lin[i]; }
    for (size_t i = 0; i < wts.size(); ++i) w_f[i] = (float)wts[i];
    std::vector<double> hw(std::max(DL, 1), 0.0);
    for (int i = 0; i < DL; ++i) { hw[i] = d->head_w[i]; hw_f[i] = (float)d->head_w[i]; }
    UP((const int4 *)d->window, (size_t)m.W0, &m.win);
    UP(slots.data(), slots.size(), &m.ent_slot);
    UP(ent_col_s.data(), (size_t)m.J0, &m.ent_col);
    UP(ent_coef_s.data(), (size_t)m.J0, &m.ent_coef);
    std::vector<double> ent_q(std::max(m.J0, 1), 0.0);
    for (int e = 0; e < m.J0; ++e) ent_q[e] = ent_coef_s[e] / d->sd[ent_col_s[e]];
    UP(ent_q.data(), ent_q.size(), &m.ent_q);
    std::vector<int32_t> sc(std::max(m.P, 1), 0);
    for (int p = 0; p < m.P; ++p) sc[p] = d->shell_count[p];
    UP(sc.data(), sc.size(), &m.shell_count);
    UP(part_shell.data(), part_shell.size(), &m.part_shell);
    UP(part_col.data(), part_col.size(), &m.part_col);
    if (n_part > 0) UP((const int4 *)d->partners, (size_t)n_part, &m.part);
    UP(c0_ch.data(), c0_ch.size(), &m.c0_ch);
    UP(d->mu, (size_t)m.C0, &m.mu); UP(d->sd, (size_t)m.C0, &m.sd); UP(d->lin, (size_t)m.C0, &m.lin);
    UP(wts.data(), wts.size(), &m.wts); UP(hw.data(), hw.size(), &m.head_w);
    UP(mu_f.data(), mu_f.size(), &m.mu_f); UP(sd_f.data(), sd_f.size(), &m.sd_f); UP(lin_f.data(), lin_f.size(), &m.lin_f);
    UP(w_f.data(), w_f.size(), &m.wts_f); UP(hw_f.data(), hw_f.size(), &m.head_w_f);
    if (m.fused_ok) {
        std::vector<uint16_t> img;
        for (int l = 0; l < m.L; ++l) {
            m.wf_off[l] = (int64_t)img.size();
            nn_stack_bf16(m, l, w_f.data() + m.off_wcat[l], c0_ch.data(), img);
        }
        UP(img.data(), img.size(), &m.wf_img);
    }
    if (pr_max >= ((int64_t)1 << 31)) return set_err(NCE_ERR_UNSUPPORTED, "projection block too large");
    m.pr_elems = pr_max;
    m.hit_elems = hit_max;
    m.w_elems = 1;
    m.mfma_ok = m.L > 0;
    for (int l = 0; l < m.L; ++l) {
        const int64_t cin = l == 0 ? m.c0_eff : m.cin[l];
        m.w_elems = std::max<int64_t>(m.w_elems, cin * m.mcat[l]);
        if (m.mcat[l] % 16 != 0) m.mfma_ok = 0;
    }
    m.lds_mode32 = nn_lds_mode<float>(m);
    m.lds_mode64 = nn_lds_mode<double>(m);
    m.lds_mode16 = nn_lds_mode<float>(m, true);
    // the BF16 layers' hash table: doubled while the fused layout keeps its placement (a wave's claims and lookups
    // wait for its longest probe chain: ~2.3 probes at 56 % load against ~1.1 at 28 %), up to 4 ball_cap[L]
    while (m.hash_cap_f < 4 * m.ball_cap[m.L]) {
        NNModelDev t = m;
        t.hash_cap_f = 2 * m.hash_cap_f;
        if (nn_lds_for<float>(t, m.lds_mode16, true).total > 160 * 1024) break;
        m.hash_cap_f = t.hash_cap_f;
    }
    m.fp32_in_lds = m.lds_mode32 == 3;
    if (m.fused_ok && nn_lds_for<float>(m, 0, true).total > 160 * 1024) m.fused_ok = 0;
    // NCE_NN_LDS_MODE_MAX=k (0..3) at model creation keeps at most the first k per-step arrays in LDS (the others in
    // the scratch): every placement computes the same step
    if (const char *cap_env = std::getenv("NCE_NN_LDS_MODE_MAX")) {
        const int cap = std::atoi(cap_env);
        if (cap < 0 || cap > 3) return set_err(NCE_ERR_ARG, "NCE_NN_LDS_MODE_MAX must be 0..3");
        m.lds_mode32 = std::min<int32_t>(m.lds_mode32, cap);
        m.lds_mode64 = std::min<int32_t>(m.lds_mode64, cap);
        m.lds_mode16 = std::min<int32_t>(m.lds_mode16, cap);
    }
    if (nn_lds_for<float>(m, 0).total > 160 * 1024 || nn_lds_for<double>(m, 0).total > 160 * 1024)
        return set_err(NCE_ERR_UNSUPPORTED, "receptive field too large for the LDS hash table");
    c->nn_models.push_back(m);
    if (int rc = nn_sync_models(c)) return rc;
    *model_id = (int32_t)c->nn_models.size() - 1;
    return NCE_OK;
}

extern "C" int nce_nn_sizes(nce_ctx *c, int32_t model, int64_t *cache_elems, int64_t *scratch_elems) {
    if (!c || model < 0 || model >= (int)c->nn_models.size()) return set_err(NCE_ERR_ARG, "unknown NeuralCE model");
    if (cache_elems) *cache_elems = c->nn_models[model].cache_elems;
    if (scratch_elems) *scratch_elems = c->nn_models[model].scratch_elems;
    return NCE_OK;
}

template <typename Real>
static int nn_full_cache_impl(nce_ctx *c, int32_t model, const int8_t *occ, int32_t R, int64_t N, Real *cache,
                              double *E, hipStream_t s) {
    const NNModelDev &m = c->nn_models[model];
    const NNModelDev *dm = (const NNModelDev *)c->d_nn_models;
    const int nb0 = (int)((N + NN_H0_SITES - 1) / NN_H0_SITES);
    const size_t lds0 = (size_t)NN_H0_SITES * (m.C0 * sizeof(Real) + nn_h0_sign_stride(m.W0, m.C0, (int)sizeof(Real)));
    if (lds0 > 160 * 1024) return set_err(NCE_ERR_UNSUPPORTED, "too many layer-0 channels");
    if (int rc = allow_lds(nn_h0_kernel<Real>, lds0)) return rc;
    const int nbe = (int)((N + 255) / 256);
    // layer-0 rows: in the cache (mode 0), else in the stream's temporary buffer, a batch of replicas at a time
    // (at most 1 GiB of rows, at least one replica)
    const int64_t row_elems = N * m.C0;
    int Rb = R;
    Real *h0 = cache;
    int64_t h0_stride = m.cache_elems;
    if (!m.cache_h0) {
        Rb = (int)std::max<int64_t>(1, std::min<int64_t>(R, ((int64_t)1 << 30) / (row_elems * (int64_t)sizeof(Real))));
        if (const char *cap = std::getenv("NCE_NN_H0_BATCH"))       // tests: smaller batches on small cells
            if (std::atoi(cap) > 0) Rb = std::min(Rb, std::atoi(cap));
        double *buf = nullptr;
        const size_t need = ((size_t)Rb * row_elems * sizeof(Real) + sizeof(double) - 1) / sizeof(double);
        if (int rc = stream_buffer(c->h0_rows, s, need, &buf)) return rc;
        h0 = (Real *)buf;
        h0_stride = row_elems;
    }
    double *partial = nullptr;
    if (int rc1 = partial_for(c, s, (size_t)Rb * nbe, &partial)) return rc1;
    for (int r0 = 0; r0 < R; r0 += Rb) {
        const int nr = std::min(Rb, R - r0);
        const int8_t *occ_b = occ + (int64_t)r0 * N;
        Real *cache_b = cache + (int64_t)r0 * m.cache_elems;
        Real *h0_b = m.cache_h0 ? cache_b : h0;
        hipLaunchKernelGGL(nn_h0_kernel<Real>, dim3(nb0, nr), dim3(NN_H0_SITES), lds0, s, dm, model, occ_b, N, h0_b,
                           h0_stride);
        HIP_TRY(hipGetLastError());
        for (int l = 0; l < m.L; ++l) {
            const int64_t nblk = ((N + 255) / 256) * ((m.dd[l] + NN_LDC - 1) / NN_LDC);
            hipLaunchKernelGGL(nn_layer_kernel<Real>, dim3((unsigned)nblk, nr), dim3(256), 0, s, dm, model, l, N, cache_b,
                               (const Real *)h0_b, h0_stride);
            HIP_TRY(hipGetLastError());
        }
        hipLaunchKernelGGL(nn_energy_kernel<Real>, dim3(nbe, nr), dim3(256), 0, s, dm, model, N, (const Real *)cache_b,
                           (const Real *)h0_b, h0_stride, partial);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(nn_energy_finish_kernel, dim3(nr), dim3(64), 0, s, dm, model, partial, nbe, E + r0);
        HIP_TRY(hipGetLastError());
    }
    return NCE_OK;
}

extern "C" int nce_nn_full_cache(nce_ctx *c, int32_t model, const int8_t *occ, int32_t R, int64_t N, void *cache,
                                 double *energy, int32_t precision, void *stream) {
    if (!c || model < 0 || model >= (int)c->nn_models.size()) return set_err(NCE_ERR_ARG, "unknown NeuralCE model");
    if (N != c->nn_models[model].n_sites) return set_err(NCE_ERR_ARG, "n_sites does not match the model's cell");
    if (R < 1 || !occ || !cache || !energy) return set_err(NCE_ERR_ARG, "bad replica count or NULL buffer");
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t s = (hipStream_t)stream;
    if (precision == NCE_PREC_FP64) return nn_full_cache_impl<double>(c, model, occ, R, N, (double *)cache, energy, s);
    if (precision == NCE_PREC_MIXED || precision == NCE_PREC_BF16)
        return nn_full_cache_impl<float>(c, model, occ, R, N, (float *)cache, energy, s);
    return set_err(NCE_ERR_ARG, "unknown precision");
}

static int nn_launch(nce_ctx *c, const nce_nn_job *jobs, int32_t n_jobs, int64_t n_steps, int32_t precision,
                     void *stream, double *eval_dE);

extern "C" int nce_nn_run_chunk(nce_ctx *c, const nce_nn_job *jobs, int32_t n_jobs, int64_t n_steps,
                                int32_t precision, void *stream) {
    return nn_launch(c, jobs, n_jobs, n_steps, precision, stream, nullptr);
}

extern "C" int nce_nn_delta(nce_ctx *c, const nce_nn_job *jobs, int32_t n_jobs, double *dE, int32_t precision,
                            void *stream) {
    if (!dE) return set_err(NCE_ERR_ARG, "NULL dE");
    for (int j = 0; j < n_jobs && jobs; ++j)
        if (!jobs[j].inj_ia) return set_err(NCE_ERR_ARG, "nce_nn_delta needs the proposals in inj_ia / inj_ib / inj_u");
    return nn_launch(c, jobs, n_jobs, 1, precision, stream, dE);
}

static int nn_launch(nce_ctx *c, const nce_nn_job *jobs, int32_t n_jobs, int64_t n_steps, int32_t precision,
                     void *stream, double *eval_dE) {
    if (!c || !jobs || n_jobs < 1) return set_err(NCE_ERR_ARG, "NULL context / jobs");
    if (n_steps < 0) return set_err(NCE_ERR_ARG, "negative n_steps");
    if (precision != NCE_PREC_FP64 && precision != NCE_PREC_MIXED && precision != NCE_PREC_BF16)
        return set_err(NCE_ERR_ARG, "unknown precision");
    if (n_steps == 0) return NCE_OK;
    HIP_TRY(hipSetDevice(c->device));
    std::vector<int2> items;
    size_t lds = 0;
    int mode = 0;
    // BF16 runs the fused BF16 layers when a model of the launch does not keep its FP32 projections in LDS (the
    // wide multi-layer networks); a launch whose every model's FP32 step fits in LDS (the shipped PtNi model) runs
    // the FP32 kernel, which is faster there (16,384 sites, 32 replicas: 2.46e6 attempts/s against 1.36e6 fused
    // over 8 workgroups).  NCE_NN_FUSED=1 forces the fused layers, 0 turns them off.
    bool mfma = precision == NCE_PREC_BF16, wide = false;
    const char *fused_env = std::getenv("NCE_NN_FUSED");
    const int fused_force = fused_env ? std::atoi(fused_env) : -1;
    for (int j = 0; j < n_jobs; ++j) {
        const nce_nn_job &jb = jobs[j];
        if (jb.model < 0 || jb.model >= (int)c->nn_models.size()) return set_err(NCE_ERR_ARG, "unknown NeuralCE model");
        const NNModelDev &m = c->nn_models[jb.model];
        if (jb.n_sites != m.n_sites || jb.n_replicas < 1) return set_err(NCE_ERR_ARG, "job sizes");
        if (!jb.occ || !jb.cache || !jb.scratch || !jb.energy || !jb.beta || !jb.step || !jb.n_accept)
            return set_err(NCE_ERR_ARG, "job has a NULL state buffer");
        if (jb.traj && (jb.sample_interval < 1 || n_steps % jb.sample_interval != 0))
            return set_err(NCE_ERR_ARG, "n_steps must be a multiple of sample_interval");
        if ((jb.inj_ia != nullptr) != (jb.inj_ib != nullptr) || (jb.inj_ia != nullptr) != (jb.inj_u != nullptr))
            return set_err(NCE_ERR_ARG, "injected streams must be all set or all NULL");
        for (int r = 0; r < jb.n_replicas; ++r) items.push_back(make_int2(j, r));
        mfma = mfma && m.fused_ok && fused_force != 0;
        wide = wide || !m.fp32_in_lds;
    }
    mfma = mfma && (fused_force == 1 || wide);
    // one LDS placement for the launch: the smallest of its models'
    for (int j = 0; j < n_jobs; ++j) {
        const NNModelDev &m = c->nn_models[jobs[j].model];
        const int mj = precision == NCE_PREC_FP64 ? m.lds_mode64 : mfma ? m.lds_mode16 : m.lds_mode32;
        mode = j == 0 ? mj : std::min(mode, mj);
    }
    for (int j = 0; j < n_jobs; ++j) {
        const NNModelDev &m = c->nn_models[jobs[j].model];
        lds = std::max(lds, precision == NCE_PREC_FP64 ? nn_lds_for<double>(m, mode).total
                                                       : nn_lds_for<float>(m, mode, mfma).total);
    }
    if (lds > 160 * 1024) return set_err(NCE_ERR_UNSUPPORTED, "receptive field does not fit in LDS");
    const size_t jb_bytes = align16(sizeof(nce_nn_job) * n_jobs);
    std::vector<unsigned char> buf(jb_bytes + items.size() * sizeof(int2));
    std::memcpy(buf.data(), jobs, sizeof(nce_nn_job) * n_jobs);
    std::memcpy(buf.data() + jb_bytes, items.data(), items.size() * sizeof(int2));
    hipStream_t s = (hipStream_t)stream;
    void *dev = nullptr;
    if (int rc = stage_upload(c, buf.data(), buf.size(), s, &dev)) return rc;
    const nce_nn_job *d_jobs = (const nce_nn_job *)dev;
    const int2 *d_items = (const int2 *)((unsigned char *)dev + jb_bytes);
    const NNModelDev *dm = (const NNModelDev *)c->d_nn_models;
    const NNSplit none{1, nullptr, nullptr};
    auto launch = [&](auto kernel) -> int {
        if (int rc = allow_lds(kernel, lds)) return rc;
        hipLaunchKernelGGL(kernel, dim3((unsigned)items.size()), dim3(NN_BS), lds, s, dm, d_jobs, d_items, n_steps,
                           eval_dE, none);
        HIP_TRY(hipGetLastError());
        return NCE_OK;
    };
    // BF16 layers of a one-job launch whose replica count is a multiple of 8: each replica over S workgroups (the
    // largest of 8, 4, 2 that fits the CUs, at most NCE_NN_SPLIT from the environment; 1 turns it off), one
    // cooperative launch so that every part is resident.  Launches of several jobs (the served and batched forms,
    // whose NeuralCE launches share the GPU with CE ticks on another stream) keep one workgroup per replica.
    if (mfma && !eval_dE && n_jobs == 1) {
        int smax = 8;
        if (const char *e = std::getenv("NCE_NN_SPLIT")) smax = std::max(1, std::atoi(e));
        const int n_items = (int)items.size();
        int S = 1;
        if (n_items % 8 == 0 && c->n_cu % 8 == 0)
            for (int cand = 8; cand >= 2; cand >>= 1)
                if (cand <= smax && n_items * cand <= c->n_cu) { S = cand; break; }
        // (only with the layer-0 change in LDS: in the scratch, one array per replica, the parts' clearing of it
        // would race with one another's adds)
        if (mode < 1) S = 1;
        int coop = 0;
        if (S > 1) HIP_TRY(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, c->device));
        if (S > 1 && coop) {
            auto kernel = mode == 0 ? nn_chunk_kernel<float, 0, true, true> : nn_chunk_kernel<float, 1, true, true>;
            if (int rc = allow_lds(kernel, lds)) return rc;
            int per_cu = 0;
            HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, NN_BS, lds));
            if (per_cu >= 1) {
                // counters (uint32, one per item) then the energy partials (double, S per item)
                const size_t ctr_d = ((size_t)n_items + 1) / 2;
                double *mem = nullptr;
                if (int rc = stream_buffer(c->split_mem, s, ctr_d + (size_t)n_items * S, &mem)) return rc;
                NNSplit sp{S, (uint32_t *)mem, mem + ctr_d};
                hipLaunchKernelGGL(nn_zero_u32, dim3((n_items + 255) / 256), dim3(256), 0, s, sp.bar, n_items);
                HIP_TRY(hipGetLastError());
                void *args[] = {(void *)&dm, (void *)&d_jobs, (void *)&d_items, (void *)&n_steps, (void *)&eval_dE,
                                (void *)&sp};
                const hipError_t e = hipLaunchCooperativeKernel((const void *)kernel, dim3((unsigned)(n_items * S)),
                                                                dim3(NN_BS), args, (unsigned)lds, s);
                if (e == hipSuccess) return NCE_OK;
                (void)hipGetLastError();               // refused (not co-resident here): one workgroup per replica
            }
        }
    }
    if (precision == NCE_PREC_FP64) {
        switch (mode) {
            case 0: return launch(nn_chunk_kernel<double, 0, false>);
            case 1: return launch(nn_chunk_kernel<double, 1, false>);
            case 2: return launch(nn_chunk_kernel<double, 2, false>);
            default: return launch(nn_chunk_kernel<double, 3, false>);
        }
    }
    if (mfma) {
        switch (mode) {
            case 0: return launch(nn_chunk_kernel<float, 0, true>);
            default: return launch(nn_chunk_kernel<float, 1, true>);
        }
    }
    switch (mode) {
        case 0: return launch(nn_chunk_kernel<float, 0, false>);
        case 1: return launch(nn_chunk_kernel<float, 1, false>);
        case 2: return launch(nn_chunk_kernel<float, 2, false>);
        default: return launch(nn_chunk_kernel<float, 3, false>);
    }
}
