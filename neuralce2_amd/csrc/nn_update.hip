// Neural cluster expansion (NeuralCE) kernels of libneuralce_hip.so, included into neuralce_hip.hip.
//
//   nn_h0_kernel          standardised layer 0 of every site from the layer-0 window and reverse entries
//   nn_layer_kernel       one hidden layer of every site (shell means of the previous layer, self + shell maps)
//   nn_energy_kernel      last layer + site energies e_i, fixed-order partial sums
//   nn_chunk_kernel       n_steps swap-Metropolis steps of one replica per workgroup with the receptive-field local
//                         update: the balls of the two flipped sites are built per step in an LDS hash table (so the
//                         two flips share one union, each site once), the layer-0 change is scattered into ball-0
//                         rows, each layer is re-evaluated on its ball from patched inputs, and on acceptance the
//                         ball rows are written back to the caches.
// Layout of the per-replica cache (elements of Real): h0s [N][c0] | hidden layer l = 1..L-1 [N][d_l] | e [N].

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
template <typename Real>
__device__ inline Real nn_sigmoid(Real x) { return (Real)1 / ((Real)1 + exp(-x)); }

template <typename Real>
__device__ inline Real nn_activate(const NNModelDev &m, Real pre, Real gpre) {
    if (m.act == 1) return pre * nn_sigmoid(pre);
    Real out = tanh(pre);
    if (m.has_gate) out *= nn_sigmoid(gpre);
    return out;
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

// layer l (0-based over hidden layers) at one site and output channel d; get(site, c) returns the input row
template <typename Real, typename Get>
__device__ inline Real nn_layer_out(const NNModelDev &m, int l, int64_t site, int d, Get get) {
    const Real *W = nn_w<Real>(m);
    const int C = m.cin[l], D = m.dd[l];
    Real pre = W[m.off_b[l] + d];
    Real gpre = m.has_gate ? W[m.off_gb[l] + d] : (Real)0;
    for (int c = 0; c < C; ++c) pre += get(site, c) * W[m.off_self[l] + (int64_t)c * D + d];
    int q0 = 0;
    for (int p = 0; p < m.P; ++p) {
        const int np_ = m.shell_count[p];
        const Real inv = np_ > 0 ? (Real)1 / (Real)np_ : (Real)0;
        for (int c = 0; c < C; ++c) {
            Real z = (Real)0;
            for (int q = 0; q < np_; ++q) z += get(nn_site(m, site, m.part[q0 + q]), c);
            z *= inv;
            pre += z * W[m.off_shell[l] + ((int64_t)p * C + c) * D + d];
            if (m.has_gate) gpre += z * W[m.off_gate[l] + ((int64_t)p * C + c) * D + d];
        }
        q0 += np_;
    }
    return nn_activate<Real>(m, pre, gpre);
}

// ── full cache ─────────────────────────────────────────────────────────────────────────────────────────────
constexpr int NN_H0_SITES = 64;

template <typename Real>
__global__ __launch_bounds__(NN_H0_SITES) void nn_h0_kernel(const NNModelDev *__restrict__ models, int model,
                                                            const int8_t *__restrict__ occ_all, int64_t N,
                                                            Real *__restrict__ cache_all) {
    extern __shared__ __align__(16) unsigned char smem[];
    const NNModelDev &m = models[model];
    Real *acc = (Real *)smem;                          // [NN_H0_SITES][C0]
    const int r = blockIdx.y;
    const int64_t site = (int64_t)blockIdx.x * NN_H0_SITES + threadIdx.x;
    const int8_t *occ = occ_all + (int64_t)r * N;
    Real *row = acc + (size_t)threadIdx.x * m.C0;
    for (int c = 0; c < m.C0; ++c) row[c] = (Real)0;
    if (site < N) {
        for (int e = 0; e < m.J0; ++e) {
            const uint2 sl = m.ent_slot[e];
            const uint32_t s4[4] = {sl.x & 0xffffu, sl.x >> 16, sl.y & 0xffffu, sl.y >> 16};
            double prod = 1.0;
            for (int q = 0; q < 4; ++q) {
                if (s4[q] == 0xffffu) continue;
                const int64_t x = nn_site(m, site, m.win[s4[q]]);
                prod *= occ[x] ? 1.0 : -1.0;
            }
            row[m.ent_col[e]] += (Real)(prod * m.ent_coef[e]);
        }
        row[0] = (Real)1;
        Real *out = cache_all + (int64_t)r * m.cache_elems + site * m.C0;
        for (int c = 0; c < m.C0; ++c) out[c] = (Real)(((double)row[c] - m.mu[c]) / m.sd[c]);
    }
}

template <typename Real>
__global__ __launch_bounds__(256) void nn_layer_kernel(const NNModelDev *__restrict__ models, int model, int l,
                                                       int64_t N, Real *__restrict__ cache_all) {
    const NNModelDev &m = models[model];
    const int r = blockIdx.y;
    const int D = m.dd[l];
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= N * D) return;
    const int64_t site = idx / D;
    const int d = (int)(idx - site * D);
    Real *cache = cache_all + (int64_t)r * m.cache_elems;
    const Real *in = cache + (l == 0 ? 0 : m.off_h[l]);
    const int C = m.cin[l];
    auto get = [&](int64_t x, int c) -> Real { return in[x * C + c]; };
    cache[m.off_h[l + 1] + site * D + d] = nn_layer_out<Real>(m, l, site, d, get);
}

// site energies (last layer evaluated here) and per-block partial sums in fixed order
template <typename Real>
__global__ __launch_bounds__(256) void nn_energy_kernel(const NNModelDev *__restrict__ models, int model, int64_t N,
                                                        Real *__restrict__ cache_all, double *__restrict__ partial) {
    __shared__ double red[4];
    const NNModelDev &m = models[model];
    const int r = blockIdx.y;
    const int64_t site = (int64_t)blockIdx.x * 256 + threadIdx.x;
    Real *cache = cache_all + (int64_t)r * m.cache_elems;
    double e = 0.0;
    if (site < N) {
        const Real *h0 = cache + site * m.C0;
        const Real *lin = nn_lin<Real>(m);
        Real acc = (Real)0;
        for (int c = 0; c < m.C0; ++c) acc += h0[c] * lin[c];
        if (m.L > 0) {
            const int l = m.L - 1, C = m.cin[l];
            const Real *in = cache + (l == 0 ? 0 : m.off_h[l]);
            auto get = [&](int64_t x, int c) -> Real { return in[x * C + c]; };
            const Real *hw = nn_headw<Real>(m);
            for (int d = 0; d < m.dd[l]; ++d) acc += nn_layer_out<Real>(m, l, site, d, get) * hw[d];
        }
        cache[m.off_e + site] = acc;
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
constexpr int NN_BS = 512;
constexpr int NN_EMPTY = -1;

struct NNLds {
    size_t keys, vals, slot_site, slot_pos, red, cnt, total;
};
__host__ __device__ inline NNLds nn_lds_layout(int hash_cap, int ball_cap) {
    NNLds L;
    size_t o = 0;
    L.keys = o;      o = align16(o + (size_t)hash_cap * sizeof(int));
    L.vals = o;      o = align16(o + (size_t)hash_cap * sizeof(int));
    L.slot_site = o; o = align16(o + (size_t)ball_cap * sizeof(int));
    L.slot_pos = o;  o = align16(o + (size_t)ball_cap * sizeof(int));
    L.red = o;       o = align16(o + (NN_BS / 64) * sizeof(double));
    L.cnt = o;       o = align16(o + 4 * sizeof(int));
    L.total = o;
    return L;
}

__device__ inline uint32_t nn_hash(int site, int mask) { return ((uint32_t)site * 2654435761u) & (uint32_t)mask; }

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

__device__ inline int nn_hash_find(const int *keys, const int *vals, int mask, int site) {
    uint32_t h = nn_hash(site, mask);
    while (true) {
        const int k = keys[h];
        if (k == site) return vals[h];
        if (k == NN_EMPTY) return -1;
        h = (h + 1) & (uint32_t)mask;
    }
}

template <typename Real>
__device__ inline Real nn_block_sum(Real v, Real *red) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    Real s = red[0];
    for (int i = 1; i < NN_BS / 64; ++i) s += red[i];
    __syncthreads();
    return s;
}

template <typename Real>
__global__ __launch_bounds__(NN_BS) void nn_chunk_kernel(const NNModelDev *__restrict__ models,
                                                         const nce_nn_job *__restrict__ jobs,
                                                         const int2 *__restrict__ items, int64_t n_steps) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int2 it = items[blockIdx.x];
    const nce_nn_job job = jobs[it.x];
    const NNModelDev &m = models[job.model];
    const int r = it.y;
    const int tid = threadIdx.x;
    const int L = m.L, C0 = m.C0;
    const int mask = m.hash_cap - 1;
    const NNLds lay = nn_lds_layout(m.hash_cap, m.ball_cap[L]);
    int *keys = (int *)(smem + lay.keys);
    int *vals = (int *)(smem + lay.vals);
    int *slot_site = (int *)(smem + lay.slot_site);
    int *slot_pos = (int *)(smem + lay.slot_pos);
    double *red = (double *)(smem + lay.red);
    int *cnt = (int *)(smem + lay.cnt);
    for (int i = tid; i < m.hash_cap; i += NN_BS) keys[i] = NN_EMPTY;
    if (tid == 0) cnt[0] = 0;
    __syncthreads();

    int8_t *occ = job.occ + (int64_t)r * job.n_sites;
    Real *cache = (Real *)job.cache + (int64_t)r * m.cache_elems;
    Real *scr = (Real *)job.scratch + (int64_t)r * m.scratch_elems;
    Real *h0new = scr + m.soff_h0;
    int *src = (int *)(scr + m.soff_src);
    Real *X = scr + m.soff_x;
    Real *enew = scr + m.soff_e;
    const Real *W = nn_w<Real>(m);
    const Real *lin = nn_lin<Real>(m);
    const Real *hw = nn_headw<Real>(m);
    double E = job.energy[r];
    const double beta = job.beta[r];
    const uint64_t step0 = (uint64_t)job.step[r];
    int64_t nacc = 0;
    const int64_t inj_base = (int64_t)r * n_steps;
    int counts[NN_MAXL + 1];

    // insert `n` candidate sites given by site_fn(k); on return the new slots are numbered and cnt[0] updated
    auto insert_sites = [&](int n, auto site_fn) {
        for (int k = tid; k < n; k += NN_BS) {
            const int site = site_fn(k);
            bool claimed;
            const int pos = nn_hash_claim(keys, mask, site, &claimed);
            if (claimed) {
                const int s = atomicAdd(&cnt[0], 1);
                vals[pos] = s;
                slot_site[s] = site;
                slot_pos[s] = pos;
            }
        }
        __syncthreads();
    };

    for (int64_t t = 0; t < n_steps; ++t) {
        int64_t ia, ib;
        double u;
        if (job.inj_ia) {
            ia = job.inj_ia[inj_base + t]; ib = job.inj_ib[inj_base + t]; u = job.inj_u[inj_base + t];
        } else {
            const Proposal p = draw_proposal_1b(job.seed, step0 + (uint64_t)t, (uint32_t)r, (uint64_t)m.n_cells);
            ia = p.ia; ib = p.ib; u = p.u;
        }
        const int si = occ[ia], sj = occ[ib];
        if (ia == ib || si == sj) {                      // no-op proposal: dE = 0, accepted (reference semantics)
            ++nacc;
        } else {
            // ball 0: the two layer-0 windows
            insert_sites(2 * m.W0, [&](int k) {
                return (int)nn_site(m, k < m.W0 ? ia : ib, m.win[k < m.W0 ? k : k - m.W0]);
            });
            counts[0] = cnt[0];
            // h0new = cached standardised rows of ball 0, then the layer-0 change of flip 1 and of flip 2
            for (int k = tid; k < counts[0] * C0; k += NN_BS) {
                const int s = k / C0, c = k - s * C0;
                h0new[k] = cache[(int64_t)slot_site[s] * C0 + c];
            }
            __syncthreads();
            for (int f = 0; f < 2; ++f) {
                const int64_t fs = f == 0 ? ia : ib;
                for (int e = tid; e < m.J0; e += NN_BS) {
                    const uint2 sl = m.ent_slot[e];
                    const uint32_t s4[4] = {sl.x & 0xffffu, sl.x >> 16, sl.y & 0xffffu, sl.y >> 16};
                    int64_t xs[4];
                    double po = 1.0, pn = 1.0;
                    for (int q = 0; q < 4; ++q) {
                        if (s4[q] == 0xffffu) { xs[q] = -1; continue; }
                        const int64_t x = nn_site(m, fs, m.win[s4[q]]);
                        xs[q] = x;
                        int sp = occ[x];
                        if (f == 1 && x == ia) sp = sj;           // flip 2 sees flip 1 applied
                        int spn = sp;
                        if (x == fs) spn = f == 0 ? sj : si;
                        po *= sp ? 1.0 : -1.0;
                        pn *= spn ? 1.0 : -1.0;
                    }
                    const double dv = (pn - po) * m.ent_coef[e];
                    if (dv != 0.0) {
                        const int col = m.ent_col[e];
                        const Real v = (Real)(dv / m.sd[col]);
                        for (int q = 0; q < 4; ++q) {
                            if (xs[q] < 0) continue;
                            const int s = nn_hash_find(keys, vals, mask, (int)xs[q]);
                            atomicAdd(&h0new[(int64_t)s * C0 + col], v);
                        }
                    }
                }
            }
            __syncthreads();
            // hidden layers on the growing balls
            for (int l = 0; l < L; ++l) {
                const int Cin = m.cin[l], D = m.dd[l];
                const int prev = counts[l];
                insert_sites(prev * m.n_part, [&](int k) {
                    const int s = k / m.n_part, q = k - s * m.n_part;
                    return (int)nn_site(m, slot_site[s], m.part[q]);
                });
                counts[l + 1] = cnt[0];
                const int nb = counts[l + 1];
                const Real *pnew = l == 0 ? h0new : scr + m.soff_h[l - 1];
                const Real *pcache = cache + (l == 0 ? 0 : m.off_h[l]);
                // source of every (ball site, partner) input: slot of the previous ball or cache row (-1 - site)
                for (int k = tid; k < nb * (1 + m.n_part); k += NN_BS) {
                    const int s = k / (1 + m.n_part), q = k - s * (1 + m.n_part);
                    const int x = q == 0 ? slot_site[s] : (int)nn_site(m, slot_site[s], m.part[q - 1]);
                    const int sx = nn_hash_find(keys, vals, mask, x);
                    src[k] = (sx >= 0 && sx < prev) ? sx : -1 - x;
                }
                __syncthreads();
                // inputs X[s] = [h(s) | mean over shell p of h(partners)] for every ball site
                const int K = Cin * (1 + m.P);
                for (int k = tid; k < nb * Cin; k += NN_BS) {
                    const int s = k / Cin, c = k - s * Cin;
                    const int *sr = src + (int64_t)s * (1 + m.n_part);
                    auto val = [&](int code) -> Real {
                        return code >= 0 ? pnew[(int64_t)code * Cin + c] : pcache[(int64_t)(-1 - code) * Cin + c];
                    };
                    X[(int64_t)s * K + c] = val(sr[0]);
                    int q0 = 0;
                    for (int p = 0; p < m.P; ++p) {
                        const int np_ = m.shell_count[p];
                        Real z = (Real)0;
                        for (int q = 0; q < np_; ++q) z += val(sr[1 + q0 + q]);
                        X[(int64_t)s * K + Cin * (1 + p) + c] = np_ > 0 ? z / (Real)np_ : (Real)0;
                        q0 += np_;
                    }
                }
                __syncthreads();
                Real *hnew = scr + m.soff_h[l];
                for (int k = tid; k < nb * D; k += NN_BS) {
                    const int s = k / D, d = k - s * D;
                    const Real *x = X + (int64_t)s * K;
                    Real pre = W[m.off_b[l] + d];
                    Real gpre = m.has_gate ? W[m.off_gb[l] + d] : (Real)0;
                    for (int c = 0; c < Cin; ++c) pre += x[c] * W[m.off_self[l] + (int64_t)c * D + d];
                    for (int p = 0; p < m.P; ++p) {
                        const Real *z = x + Cin * (1 + p);
                        for (int c = 0; c < Cin; ++c) {
                            pre += z[c] * W[m.off_shell[l] + ((int64_t)p * Cin + c) * D + d];
                            if (m.has_gate) gpre += z[c] * W[m.off_gate[l] + ((int64_t)p * Cin + c) * D + d];
                        }
                    }
                    hnew[k] = nn_activate<Real>(m, pre, gpre);
                }
                __syncthreads();
            }
            // site energies on the last ball, energy change
            const int nL = counts[L];
            const Real *hL = L > 0 ? scr + m.soff_h[L - 1] : nullptr;
            const int DL = L > 0 ? m.dd[L - 1] : 0;
            double part = 0.0;
            for (int s = tid; s < nL; s += NN_BS) {
                const int x = slot_site[s];
                const Real *h0 = s < counts[0] ? h0new + (int64_t)s * C0 : cache + (int64_t)x * C0;
                Real e = (Real)0;
                for (int c = 0; c < C0; ++c) e += h0[c] * lin[c];
                for (int d = 0; d < DL; ++d) e += hL[(int64_t)s * DL + d] * hw[d];
                enew[s] = e;
                part += (double)e - (double)cache[m.off_e + x];
            }
            const double diff = nn_block_sum<double>(part, red);
            const double dE = (m.readout == 0 ? diff / (double)m.nA : diff);
            const double dE_eV = dE * m.de_scale;
            const bool accept = (dE_eV <= 0.0) || (u < exp(-beta * dE_eV));
            if (accept) {
                for (int k = tid; k < counts[0] * C0; k += NN_BS) {
                    const int s = k / C0, c = k - s * C0;
                    cache[(int64_t)slot_site[s] * C0 + c] = h0new[k];
                }
                for (int l = 0; l + 1 < L; ++l) {
                    const int D = m.dd[l];
                    const Real *hnew = scr + m.soff_h[l];
                    for (int k = tid; k < counts[l + 1] * D; k += NN_BS) {
                        const int s = k / D, d = k - s * D;
                        cache[m.off_h[l + 1] + (int64_t)slot_site[s] * D + d] = hnew[k];
                    }
                }
                for (int s = tid; s < nL; s += NN_BS) cache[m.off_e + slot_site[s]] = enew[s];
                if (tid == 0) {
                    occ[ia] = (int8_t)sj;
                    occ[ib] = (int8_t)si;
                }
                E += dE;
                ++nacc;
            }
            // clear the table for the next step
            for (int s = tid; s < nL; s += NN_BS) keys[slot_pos[s]] = NN_EMPTY;
            __syncthreads();
            if (tid == 0) cnt[0] = 0;
        }
        if (job.traj && ((t + 1) % job.sample_interval) == 0 && tid == 0)
            job.traj[((t + 1) / job.sample_interval - 1) * job.n_replicas + r] = E;
        __syncthreads();
    }
    if (tid == 0) {
        job.energy[r] = E;
        job.n_accept[r] += nacc;
        job.step[r] = (int64_t)(step0 + (uint64_t)n_steps);
    }
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

extern "C" int nce_nn_model_create(nce_ctx *c, const nce_nn_model_desc *d, int32_t *model_id) {
    if (!c || !d || !model_id) return set_err(NCE_ERR_ARG, "NULL argument");
    HIP_TRY(hipSetDevice(c->device));
    if (d->n_layers < 0 || d->n_layers > NN_MAXL - 1) return set_err(NCE_ERR_ARG, "0..7 hidden layers supported");
    if (d->n_window < 1 || d->n_window >= 0xFFFF || d->n_entries < 1 || d->c0 < 1 || d->n_shells < 0)
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
    m.head_b = d->head_b; m.de_scale = d->de_scale;
    int n_part = 0;
    for (int p = 0; p < m.P; ++p) {
        if (d->shell_count[p] < 0) return set_err(NCE_ERR_ARG, "negative shell count");
        n_part += d->shell_count[p];
    }
    m.n_part = n_part;
    // weights: per layer self [cin][d], shell [P][cin][d], gate [P][cin][d], b [d], gb [d]
    std::vector<double> wts;
    int64_t so = 0, sho = 0, go = 0, bo = 0, gbo = 0;
    int cin_prev = m.C0, kmax = 0;
    for (int l = 0; l < m.L; ++l) {
        const int C = d->layer_cin[l], D = d->layer_d[l];
        if (C != cin_prev || D < 1) return set_err(NCE_ERR_ARG, "layer widths do not chain");
        m.cin[l] = C; m.dd[l] = D;
        kmax = std::max(kmax, C * (1 + m.P));
        m.off_self[l] = (int64_t)wts.size();
        wts.insert(wts.end(), d->w_self + so, d->w_self + so + (int64_t)C * D); so += (int64_t)C * D;
        m.off_shell[l] = (int64_t)wts.size();
        wts.insert(wts.end(), d->w_shell + sho, d->w_shell + sho + (int64_t)m.P * C * D); sho += (int64_t)m.P * C * D;
        m.off_gate[l] = (int64_t)wts.size();
        if (m.has_gate) { wts.insert(wts.end(), d->w_gate + go, d->w_gate + go + (int64_t)m.P * C * D); go += (int64_t)m.P * C * D; }
        m.off_b[l] = (int64_t)wts.size();
        wts.insert(wts.end(), d->b + bo, d->b + bo + D); bo += D;
        m.off_gb[l] = (int64_t)wts.size();
        if (m.has_gate) { wts.insert(wts.end(), d->gb + gbo, d->gb + gbo + D); gbo += D; }
        cin_prev = D;
    }
    if (wts.empty()) wts.push_back(0.0);
    m.kmax = kmax;
    const int DL = m.L > 0 ? m.dd[m.L - 1] : 0;
    // slots
    std::vector<uint2> slots(m.J0);
    for (int e = 0; e < m.J0; ++e) {
        uint32_t q[4];
        for (int k = 0; k < 4; ++k) {
            q[k] = d->entry_slot[(size_t)e * 4 + k];
            if (q[k] != 0xFFFFu && (int)q[k] >= m.W0) return set_err(NCE_ERR_ARG, "entry slot outside the window");
        }
        if (d->entry_col[e] < 1 || d->entry_col[e] >= m.C0) return set_err(NCE_ERR_ARG, "entry channel out of range");
        slots[e] = make_uint2(q[0] | (q[1] << 16), q[2] | (q[3] << 16));
    }
    auto chk4 = [&](const int32_t *w, int n) {
        for (int i = 0; i < n; ++i)
            if (w[4 * i] != 0 || w[4 * i + 1] < 0 || w[4 * i + 1] >= m.d0 || w[4 * i + 2] < 0 || w[4 * i + 2] >= m.d1 ||
                w[4 * i + 3] < 0 || w[4 * i + 3] >= m.d2)
                return false;
        return true;
    };
    if (!chk4(d->window, m.W0) || !chk4(d->partners, n_part)) return set_err(NCE_ERR_ARG, "offsets not reduced");
    std::vector<int32_t> part_shell(std::max(n_part, 1));
    for (int p = 0, k = 0; p < m.P; ++p)
        for (int q = 0; q < d->shell_count[p]; ++q) part_shell[k++] = p;
    for (int l = 0; l <= m.L; ++l) {
        if (d->ball_cap[l] < 1) return set_err(NCE_ERR_ARG, "ball capacity");
        m.ball_cap[l] = d->ball_cap[l];
    }
    int hc = 1;
    while (hc < 2 * m.ball_cap[m.L]) hc <<= 1;
    m.hash_cap = hc;
    // per-replica cache layout
    const int64_t N = m.n_sites;
    int64_t o = N * m.C0;
    for (int l = 0; l + 1 < m.L; ++l) { m.off_h[l + 1] = o; o += N * m.dd[l]; }
    m.off_e = o; o += N;
    m.cache_elems = o;
    // per-replica scratch layout
    int64_t so2 = 0;
    m.soff_h0 = so2; so2 += (int64_t)m.ball_cap[0] * m.C0;
    m.soff_src = so2; so2 += (int64_t)m.ball_cap[m.L] * (1 + n_part);     // one int per Real-sized element
    m.soff_x = so2; so2 += (int64_t)m.ball_cap[m.L] * std::max(kmax, 1);
    for (int l = 0; l < m.L; ++l) { m.soff_h[l] = so2; so2 += (int64_t)m.ball_cap[l + 1] * m.dd[l]; }
    m.soff_e = so2; so2 += m.ball_cap[m.L];
    m.scratch_elems = so2;
    (void)DL;
    std::vector<float> mu_f(m.C0), sd_f(m.C0), lin_f(m.C0), w_f(wts.size()), hw_f(std::max(DL, 1));
    for (int i = 0; i < m.C0; ++i) { mu_f[i] = (float)d->mu[i]; sd_f[i] = (float)d->sd[i]; lin_f[i] = (float)d->lin[i]; }
    for (size_t i = 0; i < wts.size(); ++i) w_f[i] = (float)wts[i];
    std::vector<double> hw(std::max(DL, 1), 0.0);
    for (int i = 0; i < DL; ++i) { hw[i] = d->head_w[i]; hw_f[i] = (float)d->head_w[i]; }
    UP((const int4 *)d->window, (size_t)m.W0, &m.win);
    UP(slots.data(), slots.size(), &m.ent_slot);
    UP(d->entry_col, (size_t)m.J0, &m.ent_col);
    UP(d->entry_coef, (size_t)m.J0, &m.ent_coef);
    std::vector<int32_t> sc(std::max(m.P, 1), 0);
    for (int p = 0; p < m.P; ++p) sc[p] = d->shell_count[p];
    UP(sc.data(), sc.size(), &m.shell_count);
    UP(part_shell.data(), part_shell.size(), &m.part_shell);
    if (n_part > 0) UP((const int4 *)d->partners, (size_t)n_part, &m.part);
    UP(d->mu, (size_t)m.C0, &m.mu); UP(d->sd, (size_t)m.C0, &m.sd); UP(d->lin, (size_t)m.C0, &m.lin);
    UP(wts.data(), wts.size(), &m.wts); UP(hw.data(), hw.size(), &m.head_w);
    UP(mu_f.data(), mu_f.size(), &m.mu_f); UP(sd_f.data(), sd_f.size(), &m.sd_f); UP(lin_f.data(), lin_f.size(), &m.lin_f);
    UP(w_f.data(), w_f.size(), &m.wts_f); UP(hw_f.data(), hw_f.size(), &m.head_w_f);
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
    const size_t lds0 = (size_t)NN_H0_SITES * m.C0 * sizeof(Real);
    if (lds0 > 160 * 1024) return set_err(NCE_ERR_UNSUPPORTED, "too many layer-0 channels");
    if (int rc = allow_lds(nn_h0_kernel<Real>, lds0)) return rc;
    hipLaunchKernelGGL(nn_h0_kernel<Real>, dim3(nb0, R), dim3(NN_H0_SITES), lds0, s, dm, model, occ, N, cache);
    HIP_TRY(hipGetLastError());
    for (int l = 0; l + 1 < m.L; ++l) {
        const int64_t n = N * m.dd[l];
        hipLaunchKernelGGL(nn_layer_kernel<Real>, dim3((unsigned)((n + 255) / 256), R), dim3(256), 0, s, dm, model,
                           l, N, cache);
        HIP_TRY(hipGetLastError());
    }
    const int nbe = (int)((N + 255) / 256);
    const size_t need = (size_t)R * nbe;
    if (need > c->partial_cap) {
        if (c->d_partial) {
            HIP_TRY(hipDeviceSynchronize());
            HIP_TRY(hipFree(c->d_partial));
        }
        HIP_TRY(hipMalloc(&c->d_partial, need * sizeof(double)));
        c->partial_cap = need;
    }
    hipLaunchKernelGGL(nn_energy_kernel<Real>, dim3(nbe, R), dim3(256), 0, s, dm, model, N, cache, c->d_partial);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(nn_energy_finish_kernel, dim3(R), dim3(64), 0, s, dm, model, c->d_partial, nbe, E);
    HIP_TRY(hipGetLastError());
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
    if (precision == NCE_PREC_MIXED) return nn_full_cache_impl<float>(c, model, occ, R, N, (float *)cache, energy, s);
    return set_err(NCE_ERR_ARG, "unknown precision");
}

extern "C" int nce_nn_run_chunk(nce_ctx *c, const nce_nn_job *jobs, int32_t n_jobs, int64_t n_steps,
                                int32_t precision, void *stream) {
    if (!c || !jobs || n_jobs < 1) return set_err(NCE_ERR_ARG, "NULL context / jobs");
    if (n_steps < 0) return set_err(NCE_ERR_ARG, "negative n_steps");
    if (precision != NCE_PREC_FP64 && precision != NCE_PREC_MIXED) return set_err(NCE_ERR_ARG, "unknown precision");
    if (n_steps == 0) return NCE_OK;
    HIP_TRY(hipSetDevice(c->device));
    std::vector<int2> items;
    size_t lds = 0;
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
        lds = std::max(lds, nn_lds_layout(m.hash_cap, m.ball_cap[m.L]).total);
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
    if (precision == NCE_PREC_FP64) {
        if (int rc = allow_lds(nn_chunk_kernel<double>, lds)) return rc;
        hipLaunchKernelGGL(nn_chunk_kernel<double>, dim3((unsigned)items.size()), dim3(NN_BS), lds, s, dm, d_jobs,
                           d_items, n_steps);
    } else {
        if (int rc = allow_lds(nn_chunk_kernel<float>, lds)) return rc;
        hipLaunchKernelGGL(nn_chunk_kernel<float>, dim3((unsigned)items.size()), dim3(NN_BS), lds, s, dm, d_jobs,
                           d_items, n_steps);
    }
    HIP_TRY(hipGetLastError());
    return NCE_OK;
}
