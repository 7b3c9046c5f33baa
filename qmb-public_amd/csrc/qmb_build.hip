// qmb_build.hip -- device-side preparation of term tables and config-set indices.
//
// Term table -> compiled records grouped into buckets by required-occupied sites (so a config
// only visits terms that can succeed on it), plus the diagonal-term subset.
// Config set -> padded 64-bit key words and an open-addressing hash table (key -> row), which
// replaces the reference's per-call lexicographic sort + binary search
// (qmb/_hamiltonian_cuda.cu:241-251, :113-131) for membership lookups.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "qmb_internal.hpp"

namespace qmb {

namespace {

constexpr int kBlock = 256;

inline unsigned blocks_for(int64_t n, int block = kBlock) {
    int64_t b = (n + block - 1) / block;
    return static_cast<unsigned>(b < 1 ? 1 : b);
}

__global__ void fill_u64_kernel(uint64_t* __restrict__ p, uint64_t value, int64_t n) {
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        p[i] = value;
    }
}

__global__ void zero_f64_kernel(double* __restrict__ p, int64_t n) {
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        p[i] = 0.0;
    }
}

// -------------------------------------------------------------------------------- terms ------
__global__ void compile_terms_kernel(const int16_t* __restrict__ site, const uint8_t* __restrict__ kind, int64_t n_terms,
                                     int particle_cut, int n_bits, int64_t n_buckets, int exclude_pairs,
                                     uint64_t* __restrict__ rec_tmp, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                     uint32_t* __restrict__ dflag, int64_t* __restrict__ counts) {
    int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= n_terms) {
        return;
    }
    CompiledTerm c = compile_term(site + 4 * t, kind + 4 * t, particle_cut, n_bits);
    vals[t] = static_cast<uint32_t>(t);
    PairExcitation pe;
    SingleExcitation se;
    if (c.status == kCompileOk && (((exclude_pairs & kExcludePairs) && pair_excitation(c.rec, particle_cut, &pe)) ||
                                   ((exclude_pairs & kExcludeSingles) && single_excitation(c.rec, particle_cut, &se)) ||
                                   ((exclude_pairs & kExcludeDiagonal) && c.diagonal) ||
                                   ((exclude_pairs & kExcludeHops) && single_excitation(c.rec, particle_cut, &se) &&
                                    ((se.r ^ se.p) & 1) == 0))) {
        // carried by the pair / single excitation tables or the diagonal kernel instead
        keys[t] = static_cast<uint32_t>(n_buckets);
        rec_tmp[t] = kNeverSucceeds;
        dflag[t] = 0u;
    } else if (c.status == kCompileOk) {
        keys[t] = static_cast<uint32_t>(bucket_of(c, n_bits));
        rec_tmp[t] = c.rec;
        dflag[t] = c.diagonal ? 1u : 0u;
    } else {
        keys[t] = static_cast<uint32_t>(n_buckets);  // sorts after every real bucket
        rec_tmp[t] = kNeverSucceeds;
        dflag[t] = 0u;
        if (c.status == kCompileBadSite || c.status == kCompileBadKind) {
            atomicAdd(reinterpret_cast<unsigned long long*>(&counts[2]), 1ull);
        }
    }
}

__global__ void gather_terms_kernel(const uint32_t* __restrict__ order, int64_t n_terms, const uint64_t* __restrict__ rec_tmp,
                                    const double2* __restrict__ coef, uint64_t* __restrict__ recs, double2* __restrict__ coefs,
                                    uint32_t* __restrict__ tidx) {
    int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (p >= n_terms) {
        return;
    }
    uint32_t t = order[p];
    recs[p] = rec_tmp[t];
    coefs[p] = coef[t];
    tidx[p] = t;
}

// bucket_start[b] = first sorted position whose key >= b, for b in [0, n_buckets].
__global__ void bucket_start_kernel(const uint32_t* __restrict__ skeys, int64_t n_terms, int64_t n_buckets,
                                    uint32_t* __restrict__ bucket_start, int64_t* __restrict__ counts) {
    int64_t b = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (b > n_buckets) {
        return;
    }
    int64_t lo = 0, hi = n_terms;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (skeys[mid] < b) {
            lo = mid + 1;
        } else {
            hi = mid;
        }
    }
    bucket_start[b] = static_cast<uint32_t>(lo);
    if (b == n_buckets) {
        counts[0] = lo;
    }
}

__global__ void scatter_diag_kernel(const uint32_t* __restrict__ dflag, const uint32_t* __restrict__ dpos, int64_t n_terms,
                                    const uint64_t* __restrict__ rec_tmp, const double2* __restrict__ coef,
                                    uint64_t* __restrict__ drecs, double2* __restrict__ dcoefs, int64_t* __restrict__ counts) {
    int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= n_terms) {
        return;
    }
    uint32_t pos = dpos[t];
    if (dflag[t]) {
        drecs[pos] = rec_tmp[t];
        dcoefs[pos] = coef[t];
    }
    if (t == n_terms - 1) {
        counts[1] = static_cast<int64_t>(pos) + dflag[t];
    }
}

__global__ void init_counts_kernel(int64_t* counts) {
    if (threadIdx.x < 4) {
        counts[threadIdx.x] = 0;
    }
}

__global__ void orig_records_kernel(const int16_t* __restrict__ site, const uint8_t* __restrict__ kind, int64_t n_terms,
                                    int particle_cut, int n_bits, uint64_t* __restrict__ orecs) {
    int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= n_terms) {
        return;
    }
    CompiledTerm c = compile_term(site + 4 * t, kind + 4 * t, particle_cut, n_bits);
    orecs[t] = c.status == kCompileOk ? c.rec : kNeverSucceeds;
}

// ------------------------------------------------------------------------------ targets ------
template <int W>
__global__ void keys_kernel(const uint8_t* __restrict__ configs, int64_t n, int nq, uint64_t* __restrict__ keys) {
    int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) {
        return;
    }
    const uint8_t* row = configs + i * nq;
    uint64_t w[W];
#pragma unroll
    for (int k = 0; k < W; ++k) {
        w[k] = 0;
    }
    for (int b = 0; b < nq; ++b) {
        const uint64_t byte = row[b];
#pragma unroll
        for (int k = 0; k < W; ++k) {
            if (k == (b >> 3)) {
                w[k] |= byte << (8 * (b & 7));
            }
        }
    }
#pragma unroll
    for (int k = 0; k < W; ++k) {
        keys[i * W + k] = w[k];
    }
}

__global__ void hash_columns_kernel(uint64_t* __restrict__ cols, int n) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < n) {
        cols[b] = hash_column(b);
    }
}

template <int W>
__global__ void insert_kernel(const uint64_t* __restrict__ keys, int64_t n, uint64_t* __restrict__ table, uint64_t mask,
                              uint64_t* __restrict__ filter, uint64_t fmask, const uint64_t* __restrict__ cols,
                              unsigned long long* __restrict__ dups) {
    int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) {
        return;
    }
    Cfg<W> c;
#pragma unroll
    for (int k = 0; k < W; ++k) {
        c.w[k] = keys[i * W + k];
    }
    const uint64_t h = hash_cfg<W>(c, cols);
    const uint32_t fp = static_cast<uint32_t>(h);
    const uint64_t entry = (static_cast<uint64_t>(i) << 32) | fp;
    uint64_t slot = (h >> 32) & mask;
    atomicOr(reinterpret_cast<unsigned long long*>(&filter[filter_word(h, fmask)]), static_cast<unsigned long long>(filter_bits(h)));
    // Bounded by the capacity: the table is at most a quarter full, so an empty slot always exists.
    for (uint64_t probe = 0; probe <= mask; ++probe) {
        unsigned long long prev = atomicCAS(reinterpret_cast<unsigned long long*>(&table[slot]),
                                            static_cast<unsigned long long>(kEmptySlot), static_cast<unsigned long long>(entry));
        if (prev == kEmptySlot) {
            return;
        }
        if (static_cast<uint32_t>(prev) == fp) {
            const uint64_t j = prev >> 32;
            Cfg<W> o;
#pragma unroll
            for (int k = 0; k < W; ++k) {
                o.w[k] = keys[j * W + k];
            }
            if (cfg_equal<W>(o, c)) {
                // duplicate row: the lowest row index keeps the slot
                atomicMin(reinterpret_cast<unsigned long long*>(&table[slot]), static_cast<unsigned long long>(entry));
                atomicAdd(dups, 1ull);
                return;
            }
        }
        slot = (slot + 1) & mask;
    }
}

// ------------------------------------------------------------------------- config sort ------
template <int W>
__global__ void be_word_kernel(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ perm, int64_t n, int word,
                               uint64_t* __restrict__ out) {
    int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (p >= n) {
        return;
    }
    out[p] = bswap64(keys[static_cast<int64_t>(perm[p]) * W + word]);
}

__global__ void iota_kernel(uint32_t* __restrict__ p, int64_t n) {
    int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i < n) {
        p[i] = static_cast<uint32_t>(i);
    }
}

template <int W>
__global__ void gather_be_kernel(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ perm, int64_t n,
                                 uint64_t* __restrict__ out) {
    int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (p >= n) {
        return;
    }
#pragma unroll
    for (int k = 0; k < W; ++k) {
        out[p * W + k] = bswap64(keys[static_cast<int64_t>(perm[p]) * W + k]);
    }
}

template <int W>
int keys_launch(const uint8_t* configs, int64_t n, int nq, uint64_t* keys, hipStream_t stream) {
    if (n > 0) {
        keys_kernel<W><<<blocks_for(n), kBlock, 0, stream>>>(configs, n, nq, keys);
        QMB_CHECK_LAUNCH();
    }
    return QMB_OK;
}

int keys_dispatch(int W, const uint8_t* configs, int64_t n, int nq, uint64_t* keys, hipStream_t stream) {
    switch (W) {
        case 1: return keys_launch<1>(configs, n, nq, keys, stream);
        case 2: return keys_launch<2>(configs, n, nq, keys, stream);
        case 3: return keys_launch<3>(configs, n, nq, keys, stream);
        default: return keys_launch<4>(configs, n, nq, keys, stream);
    }
}

template <int W>
int insert_launch(const uint64_t* keys, int64_t n, uint64_t* table, uint64_t mask, uint64_t* filter, uint64_t fmask,
                  const uint64_t* cols, int64_t* dups, hipStream_t stream) {
    if (n > 0) {
        insert_kernel<W><<<blocks_for(n), kBlock, 0, stream>>>(keys, n, table, mask, filter, fmask, cols,
                                                                 reinterpret_cast<unsigned long long*>(dups));
        QMB_CHECK_LAUNCH();
    }
    return QMB_OK;
}

int insert_dispatch(int W, const uint64_t* keys, int64_t n, uint64_t* table, uint64_t mask, uint64_t* filter, uint64_t fmask,
                    const uint64_t* cols, int64_t* dups, hipStream_t stream) {
    switch (W) {
        case 1: return insert_launch<1>(keys, n, table, mask, filter, fmask, cols, dups, stream);
        case 2: return insert_launch<2>(keys, n, table, mask, filter, fmask, cols, dups, stream);
        case 3: return insert_launch<3>(keys, n, table, mask, filter, fmask, cols, dups, stream);
        default: return insert_launch<4>(keys, n, table, mask, filter, fmask, cols, dups, stream);
    }
}

uint32_t key_bits(int64_t n_buckets) {
    uint32_t bits = 1;
    while ((1ll << bits) <= n_buckets) {
        ++bits;
    }
    return bits;
}

}  // namespace

int launch_fill_u64(uint64_t* p, uint64_t value, int64_t n, hipStream_t stream) {
    if (n > 0) {
        int64_t blocks = (n + kBlock - 1) / kBlock;
        fill_u64_kernel<<<static_cast<unsigned>(blocks < 65536 ? blocks : 65536), kBlock, 0, stream>>>(p, value, n);
        QMB_CHECK_LAUNCH();
    }
    return QMB_OK;
}

int launch_zero_f64(double* p, int64_t n, hipStream_t stream) {
    if (n > 0) {
        int64_t blocks = (n + kBlock - 1) / kBlock;
        zero_f64_kernel<<<static_cast<unsigned>(blocks < 65536 ? blocks : 65536), kBlock, 0, stream>>>(p, n);
        QMB_CHECK_LAUNCH();
    }
    return QMB_OK;
}

// ------------------------------------------------------------------------------- layouts -----
int terms_layout(Carver& cv, int n_qubytes, int64_t n_terms, qmb_terms* out, void** scratch, size_t* scratch_bytes) {
    QMB_TRY(check_qubytes(n_qubytes));
    if (n_terms < 0 || n_terms >= (1ll << 32) - 1) {
        return set_error(QMB_ERR_ARGUMENT, "n_terms out of range: %lld", static_cast<long long>(n_terms));
    }
    const int n_bits = 8 * n_qubytes;
    const int64_t nb = bucket_count(n_bits);
    const size_t T = static_cast<size_t>(n_terms > 0 ? n_terms : 1);
    qmb_terms t{};
    t.n_qubytes = n_qubytes;
    t.n_words = n_words_of(n_qubytes);
    t.n_bits = n_bits;
    t.n_terms = n_terms;
    t.n_buckets = nb;
    t.buffer = cv.base;
    t.counts = cv.take<int64_t>(4);
    t.recs = cv.take<uint64_t>(T);
    t.coefs = cv.take<double>(2 * T);
    t.tidx = cv.take<uint32_t>(T);
    t.bucket_start = cv.take<uint32_t>(static_cast<size_t>(nb) + 1);
    t.drecs = cv.take<uint64_t>(T);
    t.dcoefs = cv.take<double>(2 * T);
    // scratch used while building (dead afterwards)
    size_t sort_bytes = 0;
    {
        uint32_t* k = nullptr;
        uint32_t* v = nullptr;
        if (rocprim::radix_sort_pairs(nullptr, sort_bytes, k, k, v, v, static_cast<unsigned>(T), 0, key_bits(nb)) !=
            hipSuccess) {
            return set_error(QMB_ERR_HIP, "rocprim radix_sort_pairs size query failed");
        }
        size_t scan_bytes = 0;
        if (rocprim::exclusive_scan(nullptr, scan_bytes, k, k, 0u, static_cast<size_t>(T), rocprim::plus<uint32_t>()) !=
            hipSuccess) {
            return set_error(QMB_ERR_HIP, "rocprim exclusive_scan size query failed");
        }
        sort_bytes = sort_bytes > scan_bytes ? sort_bytes : scan_bytes;
    }
    size_t scratch_start = cv.off;
    void* s = cv.take_bytes(6 * T * sizeof(uint32_t) + T * sizeof(uint64_t) + 4096 + sort_bytes);
    if (scratch) {
        *scratch = s;
    }
    if (scratch_bytes) {
        *scratch_bytes = cv.off - scratch_start;
    }
    t.bytes = cv.off;
    if (out) {
        *out = t;
    }
    return QMB_OK;
}

int terms_build(int n_qubytes, int particle_cut, const int16_t* site, const uint8_t* kind, const double* coef, int64_t n_terms,
                void* buffer, size_t bytes, qmb_terms* out, hipStream_t stream, int exclude_pairs) {
    if (particle_cut != 1 && particle_cut != 2) {
        return set_error(QMB_ERR_UNSUPPORTED, "particle_cut must be 1 or 2, got %d", particle_cut);
    }
    Carver cv{static_cast<char*>(buffer), 0};
    qmb_terms t{};
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    QMB_TRY(terms_layout(cv, n_qubytes, n_terms, &t, &scratch, &scratch_bytes));
    if (cv.off > bytes) {
        return set_error(QMB_ERR_WORKSPACE, "term buffer too small: need %zu bytes, have %zu", cv.off, bytes);
    }
    t.particle_cut = particle_cut;
    t.bytes = bytes;
    const size_t T = static_cast<size_t>(n_terms > 0 ? n_terms : 1);
    Carver sc{static_cast<char*>(scratch), 0};
    uint32_t* keys = sc.take<uint32_t>(T);
    uint32_t* skeys = sc.take<uint32_t>(T);
    uint32_t* vals = sc.take<uint32_t>(T);
    uint32_t* order = sc.take<uint32_t>(T);
    uint32_t* dflag = sc.take<uint32_t>(T);
    uint32_t* dpos = sc.take<uint32_t>(T);
    uint64_t* rec_tmp = sc.take<uint64_t>(T);
    size_t tmp_bytes = scratch_bytes > sc.off + 512 ? scratch_bytes - sc.off - 512 : 0;
    void* tmp = sc.take_bytes(tmp_bytes);

    init_counts_kernel<<<1, 64, 0, stream>>>(t.counts);
    QMB_CHECK_LAUNCH();
    if (n_terms > 0) {
        compile_terms_kernel<<<blocks_for(n_terms), kBlock, 0, stream>>>(site, kind, n_terms, particle_cut, t.n_bits,
                                                                        t.n_buckets, exclude_pairs, rec_tmp, keys, vals, dflag,
                                                                        t.counts);
        QMB_CHECK_LAUNCH();
        size_t need = tmp_bytes;
        QMB_HIP_TRY(rocprim::radix_sort_pairs(tmp, need, keys, skeys, vals, order, static_cast<unsigned>(n_terms), 0,
                                              key_bits(t.n_buckets), stream));
        gather_terms_kernel<<<blocks_for(n_terms), kBlock, 0, stream>>>(order, n_terms, rec_tmp,
                                                                       reinterpret_cast<const double2*>(coef), t.recs,
                                                                       reinterpret_cast<double2*>(t.coefs), t.tidx);
        QMB_CHECK_LAUNCH();
        need = tmp_bytes;
        QMB_HIP_TRY(rocprim::exclusive_scan(tmp, need, dflag, dpos, 0u, static_cast<size_t>(n_terms),
                                            rocprim::plus<uint32_t>(), stream));
        scatter_diag_kernel<<<blocks_for(n_terms), kBlock, 0, stream>>>(dflag, dpos, n_terms, rec_tmp,
                                                                       reinterpret_cast<const double2*>(coef), t.drecs,
                                                                       reinterpret_cast<double2*>(t.dcoefs), t.counts);
        QMB_CHECK_LAUNCH();
    }
    bucket_start_kernel<<<blocks_for(t.n_buckets + 1), kBlock, 0, stream>>>(skeys, n_terms, t.n_buckets, t.bucket_start,
                                                                          t.counts);
    QMB_CHECK_LAUNCH();
    *out = t;
    return QMB_OK;
}

int compile_records_original(int n_qubytes, int particle_cut, const int16_t* site, const uint8_t* kind, int64_t n_terms,
                             uint64_t* orecs, hipStream_t stream) {
    if (n_terms > 0) {
        orig_records_kernel<<<blocks_for(n_terms), kBlock, 0, stream>>>(site, kind, n_terms, particle_cut, 8 * n_qubytes,
                                                                       orecs);
        QMB_CHECK_LAUNCH();
    }
    return QMB_OK;
}

int targets_layout(Carver& cv, int n_qubytes, int64_t n, qmb_targets* out) {
    QMB_TRY(check_qubytes(n_qubytes));
    if (n < 0 || n >= (1ll << 32) - 1) {
        return set_error(QMB_ERR_ARGUMENT, "config count out of range: %lld", static_cast<long long>(n));
    }
    qmb_targets t{};
    t.n_qubytes = n_qubytes;
    t.n_words = n_words_of(n_qubytes);
    t.n = n;
    t.capacity = next_pow2(static_cast<uint64_t>(4 * n > 64 ? 4 * n : 64));  // load factor <= 1/4
    t.buffer = cv.base;
    // 16 bits per key up to 2 MB of filter (~0.5% false positives), then 8 bits per key (~3%): a
    // filter that stays in an XCD's 4 MB L2 answers more of the probes than a sparser one
    // that spills (4M configs: 8 MB -> 4 MB measured 3% faster over the whole lookup)
    constexpr uint64_t kFilterWords16 = 1ull << 18;
    const uint64_t w16 = next_pow2(static_cast<uint64_t>(n / 4 > 64 ? n / 4 : 64));
    t.filter_words = w16 <= kFilterWords16 ? w16 : std::max(kFilterWords16, next_pow2(static_cast<uint64_t>(n / 8)));
    t.keys = cv.take<uint64_t>(static_cast<size_t>(n > 0 ? n : 1) * t.n_words);
    t.table = cv.take<uint64_t>(t.capacity);
    t.filter = cv.take<uint64_t>(t.filter_words);
    t.cols = cv.take<uint64_t>(64 * kMaxWords);
    t.duplicates = cv.take<int64_t>(1);
    t.bytes = cv.off;
    if (out) {
        *out = t;
    }
    return QMB_OK;
}

int targets_build_keys(int n_qubytes, const uint8_t* configs, int64_t n, void* buffer, size_t bytes, qmb_targets* out,
                       hipStream_t stream) {
    Carver cv{static_cast<char*>(buffer), 0};
    qmb_targets t{};
    QMB_TRY(targets_layout(cv, n_qubytes, n, &t));
    if (cv.off > bytes) {
        return set_error(QMB_ERR_WORKSPACE, "target buffer too small: need %zu bytes, have %zu", cv.off, bytes);
    }
    t.bytes = bytes;
    QMB_TRY(keys_dispatch(t.n_words, configs, n, n_qubytes, t.keys, stream));
    hash_columns_kernel<<<1, 64 * kMaxWords, 0, stream>>>(t.cols, 64 * kMaxWords);
    QMB_CHECK_LAUNCH();
    *out = t;
    return QMB_OK;
}

int targets_build_index(const qmb_targets* t, hipStream_t stream) {
    QMB_TRY(launch_fill_u64(t->table, kEmptySlot, static_cast<int64_t>(t->capacity), stream));
    QMB_TRY(launch_fill_u64(t->filter, 0ull, static_cast<int64_t>(t->filter_words), stream));
    QMB_TRY(launch_fill_u64(reinterpret_cast<uint64_t*>(t->duplicates), 0ull, 1, stream));
    return insert_dispatch(t->n_words, t->keys, t->n, t->table, t->capacity - 1, t->filter, t->filter_words - 1, t->cols, t->duplicates,
                           stream);
}

int targets_build(int n_qubytes, const uint8_t* configs, int64_t n, void* buffer, size_t bytes, qmb_targets* out,
                  hipStream_t stream) {
    QMB_TRY(targets_build_keys(n_qubytes, configs, n, buffer, bytes, out, stream));
    return targets_build_index(out, stream);
}

// --------------------------------------------------------------------------- config sort -----
size_t sort_configs_scratch_bytes(int n_words, int64_t n) {
    const size_t N = static_cast<size_t>(n > 0 ? n : 1);
    size_t sort_bytes = 0;
    uint64_t* k = nullptr;
    uint32_t* v = nullptr;
    (void)rocprim::radix_sort_pairs(nullptr, sort_bytes, k, k, v, v, static_cast<unsigned>(N));
    Carver cv;
    cv.take<uint64_t>(N * n_words);  // raw keys
    cv.take<uint64_t>(N);            // word in
    cv.take<uint64_t>(N);            // word out
    cv.take<uint32_t>(N);            // perm in
    cv.take_bytes(sort_bytes);
    return cv.off + 256;
}

int sort_configs(int n_qubytes, const uint8_t* configs, int64_t n, uint64_t* keys_be, uint32_t* perm, void* scratch,
                 size_t scratch_bytes, hipStream_t stream) {
    const int W = n_words_of(n_qubytes);
    if (n <= 0) {
        return QMB_OK;
    }
    const size_t N = static_cast<size_t>(n);
    Carver cv{static_cast<char*>(scratch), 0};
    uint64_t* raw = cv.take<uint64_t>(N * W);
    uint64_t* win = cv.take<uint64_t>(N);
    uint64_t* wout = cv.take<uint64_t>(N);
    uint32_t* pin = cv.take<uint32_t>(N);
    size_t sort_bytes = 0;
    (void)rocprim::radix_sort_pairs(nullptr, sort_bytes, win, wout, pin, perm, static_cast<unsigned>(N));
    void* tmp = cv.take_bytes(sort_bytes);
    if (cv.off > scratch_bytes) {
        return set_error(QMB_ERR_WORKSPACE, "sort scratch too small: need %zu, have %zu", cv.off, scratch_bytes);
    }
    QMB_TRY(keys_dispatch(W, configs, n, n_qubytes, raw, stream));
    iota_kernel<<<blocks_for(n), kBlock, 0, stream>>>(perm, n);
    QMB_CHECK_LAUNCH();
    // LSD over words: least significant word (last) first; each pass is a stable radix sort.
    for (int w = W - 1; w >= 0; --w) {
        switch (W) {
            case 1: be_word_kernel<1><<<blocks_for(n), kBlock, 0, stream>>>(raw, perm, n, w, win); break;
            case 2: be_word_kernel<2><<<blocks_for(n), kBlock, 0, stream>>>(raw, perm, n, w, win); break;
            case 3: be_word_kernel<3><<<blocks_for(n), kBlock, 0, stream>>>(raw, perm, n, w, win); break;
            default: be_word_kernel<4><<<blocks_for(n), kBlock, 0, stream>>>(raw, perm, n, w, win); break;
        }
        QMB_CHECK_LAUNCH();
        QMB_HIP_TRY(hipMemcpyAsync(pin, perm, N * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream));
        size_t need = sort_bytes;
        QMB_HIP_TRY(rocprim::radix_sort_pairs(tmp, need, win, wout, pin, perm, static_cast<unsigned>(N), 0, 64, stream));
    }
    switch (W) {
        case 1: gather_be_kernel<1><<<blocks_for(n), kBlock, 0, stream>>>(raw, perm, n, keys_be); break;
        case 2: gather_be_kernel<2><<<blocks_for(n), kBlock, 0, stream>>>(raw, perm, n, keys_be); break;
        case 3: gather_be_kernel<3><<<blocks_for(n), kBlock, 0, stream>>>(raw, perm, n, keys_be); break;
        default: gather_be_kernel<4><<<blocks_for(n), kBlock, 0, stream>>>(raw, perm, n, keys_be); break;
    }
    QMB_CHECK_LAUNCH();
    return QMB_OK;
}

}  // namespace qmb
