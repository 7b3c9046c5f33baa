// qmb_cpu.hip -- the operators on host memory for CPU tensors: apply_within, its adjoint and
// find_relative, on host threads.  No device code: this is the CPU dispatch key's implementation
// (the reference registers apply_within and find_relative for CPU, qmb/_hamiltonian_cpu.cpp:558-561).
//
// Same term semantics as the GPU path (compiled records, qmb_common.hpp) and a different search
// than the reference's serial term-outer grid with binary search (qmb/_hamiltonian_cpu.cpp:131-162):
//   - terms are bucketed by the set of sites they require occupied, so a config only walks the
//     buckets its occupied sites select (bucket_of);
//   - a target is found in an open-addressing hash index over the target / exclude rows, keyed by
//     the linear site hash (hash_cfg), which a term moves by one XOR;
//   - configs_i is cut into one contiguous slice per thread.  The forward scatter gives each thread
//     its own accumulator, summed in thread order at the end (results do not depend on timing); the
//     adjoint is a gather, each thread writing only its own rows;
//   - find_relative keeps per thread the best entries under the total order (weight descending,
//     config bytes ascending), pruned to count_selected whenever the buffer doubles, then merges.
// Weights and products are computed without FMA contraction (-ffp-contract=off for this file), as
// the GPU kernels and the oracle do.
#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

#include "qmb_internal.hpp"

namespace qmb {
namespace {

template <int W>
Cfg<W> row_words(const uint8_t* row, int nq) {
    Cfg<W> c;
    for (int w = 0; w < W; ++w) {
        c.w[w] = 0;
    }
    for (int b = 0; b < nq; ++b) {
        c.w[b >> 3] |= static_cast<uint64_t>(row[b]) << (8 * (b & 7));
    }
    return c;
}

template <int W>
void words_row(const Cfg<W>& c, int nq, uint8_t* row) {
    for (int b = 0; b < nq; ++b) {
        row[b] = static_cast<uint8_t>(c.w[b >> 3] >> (8 * (b & 7)));
    }
}

uint64_t host_cols[256];
struct ColsInit {
    ColsInit() {
        for (int s = 0; s < 256; ++s) {
            host_cols[s] = hash_column(s);
        }
    }
} cols_init;

// Term table in bucket order: the expanded masks of every term that can succeed, its coefficient
// and the bucket boundaries.
template <int W>
struct HostTerms {
    int n_bits = 0;
    std::vector<Masks<W>> masks;
    std::vector<double> coef;  // [2 k]
    std::vector<int64_t> start;  // [buckets + 1]
};

template <int W>
int build_terms(int nq, int pc, const int16_t* site, const uint8_t* kind, const double* coef, int64_t n_terms, HostTerms<W>& out) {
    const int n_bits = 8 * nq;
    const int64_t n_buckets = bucket_count(n_bits);
    std::vector<int64_t> bucket(static_cast<size_t>(n_terms), -1);
    std::vector<uint64_t> rec(static_cast<size_t>(n_terms), 0);
    out.n_bits = n_bits;
    out.start.assign(static_cast<size_t>(n_buckets + 1), 0);
    for (int64_t t = 0; t < n_terms; ++t) {
        const CompiledTerm c = compile_term(site + 4 * t, kind + 4 * t, pc, n_bits);
        if (c.status == kCompileBadSite || c.status == kCompileBadKind) {
            return set_error(QMB_ERR_ARGUMENT, "term %lld has a kind outside {0,1,2} or a site outside [0, %d)",
                             static_cast<long long>(t), n_bits);
        }
        if (c.status != kCompileOk) {
            continue;  // statically zero (e.g. a_p a_p)
        }
        bucket[t] = bucket_of(c, n_bits);
        rec[t] = c.rec;
        ++out.start[bucket[t] + 1];
    }
    for (int64_t b = 0; b < n_buckets; ++b) {
        out.start[b + 1] += out.start[b];
    }
    const int64_t n_valid = out.start[n_buckets];
    out.masks.resize(static_cast<size_t>(n_valid));
    out.coef.resize(static_cast<size_t>(2 * n_valid));
    std::vector<int64_t> fill(out.start.begin(), out.start.end() - 1);
    for (int64_t t = 0; t < n_terms; ++t) {  // stable: term order inside a bucket
        if (bucket[t] < 0) {
            continue;
        }
        const int64_t k = fill[bucket[t]]++;
        out.masks[k] = expand_rec<W>(rec[t], host_cols);
        out.coef[2 * k] = coef[2 * t];
        out.coef[2 * k + 1] = coef[2 * t + 1];
    }
    return QMB_OK;
}

// Hash index over a config set: row numbers in an open-addressing table of 2^k >= 2 n slots,
// linear probing; a repeated config keeps its lowest row.
template <int W>
struct HostIndex {
    std::vector<Cfg<W>> keys;
    std::vector<uint64_t> hashes;
    std::vector<int64_t> slots;
    uint64_t mask = 0;

    void build(const uint8_t* configs, int64_t n, int nq) {
        keys.resize(static_cast<size_t>(n));
        hashes.resize(static_cast<size_t>(n));
        for (int64_t r = 0; r < n; ++r) {
            keys[r] = row_words<W>(configs + r * nq, nq);
            hashes[r] = hash_cfg<W>(keys[r], host_cols);
        }
        const uint64_t cap = next_pow2(static_cast<uint64_t>(std::max<int64_t>(2 * n, 2)));
        mask = cap - 1;
        slots.assign(static_cast<size_t>(cap), -1);
        for (int64_t r = 0; r < n; ++r) {
            uint64_t s = fmix64(hashes[r]) & mask;
            while (slots[s] >= 0 && !cfg_equal<W>(keys[slots[s]], keys[r])) {
                s = (s + 1) & mask;
            }
            if (slots[s] < 0) {
                slots[s] = r;
            }
        }
    }

    int64_t find(const Cfg<W>& c, uint64_t h) const {
        uint64_t s = fmix64(h) & mask;
        while (true) {
            const int64_t r = slots[s];
            if (r < 0) {
                return -1;
            }
            if (hashes[r] == h && cfg_equal<W>(keys[r], c)) {
                return r;
            }
            s = (s + 1) & mask;
        }
    }
};

// Calls fn(term k, target, parity) for every term of `terms` that succeeds on c.
template <int W, typename F>
void for_each_connection(const HostTerms<W>& terms, const Cfg<W>& c, F&& fn) {
    int occ[256];
    int ne = 0;
    for (int w = 0; w < W; ++w) {
        uint64_t x = c.w[w];
        while (x) {
            occ[ne++] = 64 * w + __builtin_ctzll(x);
            x &= x - 1;
        }
    }
    const int n_bits = terms.n_bits;
    auto walk = [&](int64_t b) {
        for (int64_t k = terms.start[b]; k < terms.start[b + 1]; ++k) {
            const Masks<W>& x = terms.masks[k];
            bool ok = true;
            for (int w = 0; w < W; ++w) {
                ok = ok && ((c.w[w] & x.m[w]) == x.v[w]);
            }
            if (!ok) {
                continue;
            }
            Cfg<W> t;
            int parity = x.pconst;
            for (int w = 0; w < W; ++w) {
                t.w[w] = c.w[w] ^ x.f[w];
                parity ^= __builtin_popcountll(c.w[w] & x.p[w]) & 1;
            }
            fn(k, t, parity);
        }
    };
    walk(0);
    for (int a = 0; a < ne; ++a) {
        walk(1 + occ[a]);
    }
    for (int b = 1; b < ne; ++b) {
        for (int a = 0; a < b; ++a) {
            walk(1 + n_bits + pair_index(occ[a], occ[b]));
        }
    }
    walk(bucket_count(n_bits) - 1);
}

int thread_count(int threads, int64_t n) {
    int t = threads > 0 ? threads : static_cast<int>(std::thread::hardware_concurrency());
    t = std::max(1, t);
    // at least 64 rows per thread
    return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(t, (n + 63) / 64)));
}

template <typename F>
void parallel_slices(int threads, int64_t n, F&& fn) {
    if (threads <= 1) {
        fn(0, int64_t{0}, n);
        return;
    }
    std::vector<std::thread> pool;
    pool.reserve(static_cast<size_t>(threads - 1));
    for (int t = 1; t < threads; ++t) {
        pool.emplace_back([&fn, t, threads, n] { fn(t, n * t / threads, n * (t + 1) / threads); });
    }
    fn(0, int64_t{0}, n / threads);
    for (auto& th : pool) {
        th.join();
    }
}

template <int W>
int cpu_apply(int nq, int pc, const uint8_t* configs_i, const double* vec_in, int64_t n_i, const uint8_t* configs_j,
              const double* vec_j, int64_t n_j, const int16_t* site, const uint8_t* kind, const double* coef, int64_t n_terms,
              double* out, int mode, int threads) {
    HostTerms<W> terms;
    QMB_TRY(build_terms<W>(nq, pc, site, kind, coef, n_terms, terms));
    HostIndex<W> index;
    index.build(configs_j, n_j, nq);
    if (mode == 1) {
        // adjoint gather: out[i] = sum conj(sign coef) g_j[target]; each thread owns its rows
        const int nt = thread_count(threads, n_i);
        parallel_slices(nt, n_i, [&](int, int64_t lo, int64_t hi) {
            for (int64_t i = lo; i < hi; ++i) {
                const Cfg<W> c = row_words<W>(configs_i + i * nq, nq);
                const uint64_t h = hash_cfg<W>(c, host_cols);
                double re = 0.0, im = 0.0;
                for_each_connection<W>(terms, c, [&](int64_t k, const Cfg<W>& t, int parity) {
                    const int64_t r = index.find(t, h ^ terms.masks[k].hf);
                    if (r < 0) {
                        return;
                    }
                    const double cr = terms.coef[2 * k], ci = terms.coef[2 * k + 1];
                    const double gr = vec_j[2 * r], gi = vec_j[2 * r + 1];
                    const double pr = cr * gr + ci * gi;
                    const double pi = cr * gi - ci * gr;
                    re += parity ? -pr : pr;
                    im += parity ? -pi : pi;
                });
                out[2 * i] = re;
                out[2 * i + 1] = im;
            }
        });
        return QMB_OK;
    }
    // forward scatter into per-thread accumulators, summed in thread order; at most 8 GiB of them
    int nt = thread_count(threads, n_i);
    const double acc_bytes = 16.0 * static_cast<double>(n_j);
    while (nt > 1 && acc_bytes * (nt - 1) > 8.0 * (1ull << 30)) {
        --nt;
    }
    std::vector<std::vector<double>> acc(static_cast<size_t>(nt > 1 ? nt - 1 : 0));
    std::fill(out, out + 2 * n_j, 0.0);
    parallel_slices(nt, n_i, [&](int t, int64_t lo, int64_t hi) {
        double* dst = out;
        if (t > 0) {
            acc[t - 1].assign(static_cast<size_t>(2 * n_j), 0.0);
            dst = acc[t - 1].data();
        }
        for (int64_t i = lo; i < hi; ++i) {
            const Cfg<W> c = row_words<W>(configs_i + i * nq, nq);
            const uint64_t h = hash_cfg<W>(c, host_cols);
            const double xr = vec_in[2 * i], xi = vec_in[2 * i + 1];
            for_each_connection<W>(terms, c, [&](int64_t k, const Cfg<W>& tg, int parity) {
                const int64_t r = index.find(tg, h ^ terms.masks[k].hf);
                if (r < 0) {
                    return;
                }
                const double cr = terms.coef[2 * k], ci = terms.coef[2 * k + 1];
                const double pr = cr * xr - ci * xi;
                const double pi = cr * xi + ci * xr;
                dst[2 * r] += parity ? -pr : pr;
                dst[2 * r + 1] += parity ? -pi : pi;
            });
        }
    });
    if (!acc.empty()) {
        parallel_slices(thread_count(threads, n_j), n_j, [&](int, int64_t lo, int64_t hi) {
            for (const auto& a : acc) {
                for (int64_t k = 2 * lo; k < 2 * hi; ++k) {
                    out[k] += a[k];
                }
            }
        });
    }
    return QMB_OK;
}

// find_relative entry: the contribution's weight and its target as byte-order key words (word 0
// holds the most significant bytes).
template <int W>
struct Entry {
    double w;
    uint64_t key[W];
};

template <int W>
bool better(const Entry<W>& a, const Entry<W>& b) {
    if (a.w != b.w) {
        return a.w > b.w;
    }
    for (int w = 0; w < W; ++w) {
        if (a.key[w] != b.key[w]) {
            return a.key[w] < b.key[w];
        }
    }
    return false;
}

template <int W>
void keep_best(std::vector<Entry<W>>& v, size_t k) {
    if (v.size() <= k) {
        return;
    }
    std::nth_element(v.begin(), v.begin() + static_cast<std::ptrdiff_t>(k), v.end(), better<W>);
    v.resize(k);
}

template <int W>
int cpu_find_relative(int nq, int pc, const uint8_t* configs_i, const double* psi_i, int64_t n_i, int64_t count_selected,
                      const int16_t* site, const uint8_t* kind, const double* coef, int64_t n_terms, const uint8_t* configs_exclude,
                      int64_t n_exclude, uint8_t* out, int64_t* out_rows, int threads) {
    HostTerms<W> terms;
    QMB_TRY(build_terms<W>(nq, pc, site, kind, coef, n_terms, terms));
    HostIndex<W> exclude;
    exclude.build(configs_exclude, n_exclude, nq);
    const size_t keep = static_cast<size_t>(count_selected);
    const size_t cap = keep + std::max<size_t>(keep, size_t{1} << 16);
    const int nt = thread_count(threads, n_i);
    std::vector<std::vector<Entry<W>>> found(static_cast<size_t>(nt));
    parallel_slices(nt, n_i, [&](int t, int64_t lo, int64_t hi) {
        std::vector<Entry<W>>& buf = found[t];
        double floor = 0.0;  // weights below the best `keep` so far can never be selected
        for (int64_t i = lo; i < hi; ++i) {
            const Cfg<W> c = row_words<W>(configs_i + i * nq, nq);
            const uint64_t h = hash_cfg<W>(c, host_cols);
            const double xr = psi_i[2 * i], xi = psi_i[2 * i + 1];
            for_each_connection<W>(terms, c, [&](int64_t k, const Cfg<W>& tg, int parity) {
                const double cr = terms.coef[2 * k], ci = terms.coef[2 * k + 1];
                const double s = parity ? -1.0 : 1.0;
                const double re = s * (cr * xr - ci * xi);
                const double im = s * (cr * xi + ci * xr);
                const double w = re * re + im * im;
                if (!(w > 0.0) || w < floor) {
                    return;  // zero weights are never kept (qmb/_hamiltonian_cpu.cpp:541-542)
                }
                if (exclude.find(tg, h ^ terms.masks[k].hf) >= 0) {
                    return;
                }
                Entry<W> e;
                e.w = w;
                for (int q = 0; q < W; ++q) {
                    e.key[q] = bswap64(tg.w[q]);
                }
                buf.push_back(e);
                if (buf.size() >= cap) {
                    keep_best(buf, keep);
                    floor = std::min_element(buf.begin(), buf.end(), [](const Entry<W>& a, const Entry<W>& b) { return a.w < b.w; })->w;
                }
            });
        }
        keep_best(buf, keep);
    });
    std::vector<Entry<W>> all;
    for (auto& v : found) {
        all.insert(all.end(), v.begin(), v.end());
        std::vector<Entry<W>>().swap(v);
    }
    keep_best(all, keep);
    std::sort(all.begin(), all.end(), [](const Entry<W>& a, const Entry<W>& b) {
        for (int w = 0; w < W; ++w) {
            if (a.key[w] != b.key[w]) {
                return a.key[w] < b.key[w];
            }
        }
        return false;
    });
    int64_t rows = 0;
    for (size_t e = 0; e < all.size(); ++e) {
        if (e > 0 && std::equal(all[e].key, all[e].key + W, all[e - 1].key)) {
            continue;
        }
        Cfg<W> c;
        for (int q = 0; q < W; ++q) {
            c.w[q] = bswap64(all[e].key[q]);
        }
        words_row<W>(c, nq, out + rows * nq);
        ++rows;
    }
    *out_rows = rows;
    return QMB_OK;
}

int check_host_args(int n_qubytes, int particle_cut, int64_t n_terms, const int16_t* site, const uint8_t* kind, const double* coef) {
    QMB_TRY(check_qubytes(n_qubytes));
    if (particle_cut != 1 && particle_cut != 2) {
        return set_error(QMB_ERR_UNSUPPORTED, "particle_cut must be 1 (fermi) or 2 (bose2), got %d", particle_cut);
    }
    if (n_terms < 0) {
        return set_error(QMB_ERR_ARGUMENT, "n_terms must be non-negative");
    }
    if (n_terms > 0 && (site == nullptr || kind == nullptr || coef == nullptr)) {
        return set_error(QMB_ERR_ARGUMENT, "site, kind and coef must be non-null when there are terms");
    }
    return QMB_OK;
}

int check_rows(const void* p, int64_t n, const char* name) {
    if (n < 0) {
        return set_error(QMB_ERR_ARGUMENT, "%s size must be non-negative", name);
    }
    if (n > 0 && p == nullptr) {
        return set_error(QMB_ERR_ARGUMENT, "%s must be non-null when its size is positive", name);
    }
    return QMB_OK;
}

template <typename F>
int by_words(int n_qubytes, F&& f) {
    switch (n_words_of(n_qubytes)) {
    case 1:
        return f(std::integral_constant<int, 1>{});
    case 2:
        return f(std::integral_constant<int, 2>{});
    case 3:
        return f(std::integral_constant<int, 3>{});
    default:
        return f(std::integral_constant<int, 4>{});
    }
}

}  // namespace
}  // namespace qmb

using namespace qmb;

extern "C" int qmb_cpu_apply_within(int n_qubytes, int particle_cut, const uint8_t* configs_i, const double* psi_i, int64_t n_i,
                                    const uint8_t* configs_j, int64_t n_j, const int16_t* site, const uint8_t* kind,
                                    const double* coef, int64_t n_terms, double* psi_j, int threads) {
    clear_error();
    QMB_TRY(check_host_args(n_qubytes, particle_cut, n_terms, site, kind, coef));
    QMB_TRY(check_rows(configs_i, n_i, "configs_i"));
    QMB_TRY(check_rows(psi_i, n_i, "psi_i"));
    QMB_TRY(check_rows(configs_j, n_j, "configs_j"));
    QMB_TRY(check_rows(psi_j, n_j, "psi_j"));
    if (n_j == 0) {
        return QMB_OK;
    }
    return by_words(n_qubytes, [&](auto wc) {
        return cpu_apply<decltype(wc)::value>(n_qubytes, particle_cut, configs_i, psi_i, n_i, configs_j, nullptr, n_j, site, kind,
                                              coef, n_terms, psi_j, 0, threads);
    });
}

extern "C" int qmb_cpu_apply_within_adjoint(int n_qubytes, int particle_cut, const uint8_t* configs_i, int64_t n_i,
                                            const uint8_t* configs_j, const double* grad_j, int64_t n_j, const int16_t* site,
                                            const uint8_t* kind, const double* coef, int64_t n_terms, double* grad_i, int threads) {
    clear_error();
    QMB_TRY(check_host_args(n_qubytes, particle_cut, n_terms, site, kind, coef));
    QMB_TRY(check_rows(configs_i, n_i, "configs_i"));
    QMB_TRY(check_rows(grad_i, n_i, "grad_i"));
    QMB_TRY(check_rows(configs_j, n_j, "configs_j"));
    QMB_TRY(check_rows(grad_j, n_j, "grad_j"));
    if (n_i == 0) {
        return QMB_OK;
    }
    return by_words(n_qubytes, [&](auto wc) {
        return cpu_apply<decltype(wc)::value>(n_qubytes, particle_cut, configs_i, nullptr, n_i, configs_j, grad_j, n_j, site, kind,
                                              coef, n_terms, grad_i, 1, threads);
    });
}

extern "C" int qmb_cpu_find_relative(int n_qubytes, int particle_cut, const uint8_t* configs_i, const double* psi_i, int64_t n_i,
                                     int64_t count_selected, const int16_t* site, const uint8_t* kind, const double* coef,
                                     int64_t n_terms, const uint8_t* configs_exclude, int64_t n_exclude, uint8_t* out,
                                     int64_t* out_rows, int threads) {
    clear_error();
    QMB_TRY(check_host_args(n_qubytes, particle_cut, n_terms, site, kind, coef));
    QMB_TRY(check_rows(configs_i, n_i, "configs_i"));
    QMB_TRY(check_rows(psi_i, n_i, "psi_i"));
    QMB_TRY(check_rows(configs_exclude, n_exclude, "configs_exclude"));
    if (count_selected < 0) {
        return set_error(QMB_ERR_ARGUMENT, "count_selected must be non-negative");
    }
    if (out_rows == nullptr || (count_selected > 0 && out == nullptr)) {
        return set_error(QMB_ERR_ARGUMENT, "out and out_rows must be non-null");
    }
    *out_rows = 0;
    if (count_selected == 0 || n_i == 0) {
        return QMB_OK;
    }
    return by_words(n_qubytes, [&](auto wc) {
        return cpu_find_relative<decltype(wc)::value>(n_qubytes, particle_cut, configs_i, psi_i, n_i, count_selected, site, kind, coef,
                                                      n_terms, configs_exclude, n_exclude, out, out_rows, threads);
    });
}
