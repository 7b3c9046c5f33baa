// qmb_capi.hip -- the extern "C" boundary of libqmb_hip.so (declared in include/qmb_hip.h).
//
// One-shot entry points mirror the reference torch ops (qmb/_hamiltonian.cpp:83-89): each one lays
// the caller's workspace out into a compiled term table and a hash index over the config set it
// needs, enqueues the builders and the kernel on the caller's stream, and returns.  Nothing here
// allocates device memory or synchronises.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <utility>
#include <vector>

#include "qmb_internal.hpp"

namespace qmb {

namespace {
thread_local std::string g_error;
}

int set_error(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

void clear_error() { g_error.clear(); }

namespace {

int check_ptr(const void* p, int64_t n, const char* name) {
    if (n > 0 && p == nullptr) {
        return set_error(QMB_ERR_ARGUMENT, "%s must be non-null when its size is positive", name);
    }
    if (n < 0) {
        return set_error(QMB_ERR_ARGUMENT, "%s size must be non-negative", name);
    }
    return QMB_OK;
}

int check_particle_cut(int pc) {
    if (pc != 1 && pc != 2) {
        return set_error(QMB_ERR_UNSUPPORTED, "particle_cut must be 1 (fermi) or 2 (bose2), got %d", pc);
    }
    return QMB_OK;
}

size_t terms_bytes(int nq, int64_t n_terms) {
    Carver cv;
    qmb_terms t{};
    if (terms_layout(cv, nq, n_terms, &t, nullptr, nullptr) != QMB_OK) {
        return 0;
    }
    return cv.off + 256;
}

size_t targets_bytes(int nq, int64_t n) {
    Carver cv;
    qmb_targets t{};
    if (targets_layout(cv, nq, n, &t) != QMB_OK) {
        return 0;
    }
    return cv.off + 256;
}

// Splits a workspace into consecutive regions of the given sizes.
struct Regions {
    char* base;
    size_t off;
    size_t cap;
    void* next(size_t bytes) {
        off = (off + 255) & ~static_cast<size_t>(255);
        void* p = base + off;
        off += bytes;
        return p;
    }
};

int need_workspace(size_t need, size_t have, void* ws) {
    if (need > have || (need > 0 && ws == nullptr)) {
        return set_error(QMB_ERR_WORKSPACE, "workspace too small: need %zu bytes, have %zu", need, have);
    }
    return QMB_OK;
}

}  // namespace
}  // namespace qmb

using namespace qmb;

extern "C" {

const char* qmb_last_error(void) { return g_error.c_str(); }

int qmb_version(void) { return 10000; }

int qmb_validate_terms_host(int n_qubytes, const int16_t* site, const uint8_t* kind, int64_t n_terms, int64_t* n_bad) {
    clear_error();
    QMB_TRY(check_qubytes(n_qubytes));
    QMB_TRY(check_ptr(site, n_terms, "site"));
    QMB_TRY(check_ptr(kind, n_terms, "kind"));
    int64_t bad = 0;
    const int n_bits = 8 * n_qubytes;
    for (int64_t t = 0; t < n_terms; ++t) {
        bool ok = true;
        for (int op = 0; op < kMaxOps; ++op) {
            const int k = kind[4 * t + op];
            if (k > 2) {
                ok = false;
            } else if (k != 2 && (site[4 * t + op] < 0 || site[4 * t + op] >= n_bits)) {
                ok = false;
            }
        }
        bad += ok ? 0 : 1;
    }
    if (n_bad) {
        *n_bad = bad;
    }
    return QMB_OK;
}

int qmb_compile_term_host(const int16_t* site4, const uint8_t* kind4, int particle_cut, int n_qubytes, uint64_t* rec,
                          int32_t* status, int64_t* bucket, int32_t* diagonal) {
    clear_error();
    QMB_TRY(check_qubytes(n_qubytes));
    QMB_TRY(check_particle_cut(particle_cut));
    const CompiledTerm c = compile_term(site4, kind4, particle_cut, 8 * n_qubytes);
    if (rec) {
        *rec = c.status == kCompileOk ? c.rec : kNeverSucceeds;
    }
    if (status) {
        *status = c.status;
    }
    if (bucket) {
        *bucket = c.status == kCompileOk ? bucket_of(c, 8 * n_qubytes) : -1;
    }
    if (diagonal) {
        *diagonal = c.status == kCompileOk && c.diagonal ? 1 : 0;
    }
    return QMB_OK;
}

int qmb_apply_record_host(uint64_t rec, int n_qubytes, const uint8_t* config, uint8_t* result, int32_t* parity) {
    clear_error();
    if (check_qubytes(n_qubytes) != QMB_OK || (rec & kNeverSucceeds)) {
        return 0;
    }
    Cfg<kMaxWords> c{};
    for (int b = 0; b < n_qubytes; ++b) {
        c.w[b >> 3] |= static_cast<uint64_t>(config[b]) << (8 * (b & 7));
    }
    Cfg<kMaxWords> o{};
    int par = 0;
    const bool ok = apply_rec<kMaxWords>(rec, c, o, par);
    if (ok) {
        for (int b = 0; b < n_qubytes; ++b) {
            result[b] = static_cast<uint8_t>(o.w[b >> 3] >> (8 * (b & 7)));
        }
        if (parity) {
            *parity = par;
        }
    }
    return ok ? 1 : 0;
}

int qmb_terms_hermitian_host(int n_qubytes, int particle_cut, const int16_t* site, const uint8_t* kind, const double* coef,
                             int64_t n_terms, double tol, int32_t* hermitian) {
    clear_error();
    QMB_TRY(check_qubytes(n_qubytes));
    QMB_TRY(check_particle_cut(particle_cut));
    QMB_TRY(check_ptr(site, n_terms, "site"));
    QMB_TRY(check_ptr(kind, n_terms, "kind"));
    QMB_TRY(check_ptr(coef, n_terms, "coef"));
    if (!hermitian) {
        return set_error(QMB_ERR_ARGUMENT, "hermitian must be non-null");
    }
    // An operator term acts as: configs with (c & m) == v map to c ^ f with sign
    // (-1)^(popcount(c & p) ^ pconst).  Its adjoint maps (c & m) == v ^ f back with the same sign
    // function (p excludes the touched sites, f lies within them), so in the key (m, v, f, p) it is
    // (m, v ^ f, f, p) with the conjugate of coef * (-1)^pconst.
    using Key = std::array<uint64_t, 4 * kMaxWords>;
    std::vector<std::pair<Key, std::pair<double, double>>> entries;
    entries.reserve(static_cast<size_t>(n_terms));
    double scale = 0.0;
    for (int64_t t = 0; t < n_terms; ++t) {
        const CompiledTerm c = compile_term(site + 4 * t, kind + 4 * t, particle_cut, 8 * n_qubytes);
        if (c.status == kCompileBadSite || c.status == kCompileBadKind) {
            return set_error(QMB_ERR_ARGUMENT, "term %lld is invalid", static_cast<long long>(t));
        }
        if (c.status != kCompileOk) {
            continue;  // the zero operator
        }
        const Masks<kMaxWords> x = expand_rec<kMaxWords>(c.rec, nullptr);
        Key k{};
        for (int w = 0; w < kMaxWords; ++w) {
            k[w] = x.m[w];
            k[kMaxWords + w] = x.v[w];
            k[2 * kMaxWords + w] = x.f[w];
            k[3 * kMaxWords + w] = x.p[w];
        }
        const double s = x.pconst ? -1.0 : 1.0;
        entries.push_back({k, {s * coef[2 * t], s * coef[2 * t + 1]}});
        scale = std::max(scale, std::hypot(coef[2 * t], coef[2 * t + 1]));
    }
    std::sort(entries.begin(), entries.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    std::vector<std::pair<Key, std::pair<double, double>>> merged;
    for (const auto& e : entries) {
        if (!merged.empty() && merged.back().first == e.first) {
            merged.back().second.first += e.second.first;
            merged.back().second.second += e.second.second;
        } else {
            merged.push_back(e);
        }
    }
    const double bound = tol * (scale > 0.0 ? scale : 1.0);
    int32_t ok = 1;
    for (const auto& e : merged) {
        Key adj = e.first;
        for (int w = 0; w < kMaxWords; ++w) {
            adj[kMaxWords + w] ^= adj[2 * kMaxWords + w];
        }
        const auto it = std::lower_bound(merged.begin(), merged.end(), adj,
                                         [](const auto& a, const Key& b) { return a.first < b; });
        double re = 0.0, im = 0.0;
        if (it != merged.end() && it->first == adj) {
            re = it->second.first;
            im = it->second.second;
        }
        if (std::hypot(re - e.second.first, im + e.second.second) > bound) {
            ok = 0;
            break;
        }
    }
    *hermitian = ok;
    return QMB_OK;
}

void qmb_philox4x32_10_host(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    U32x4 c{{ctr[0], ctr[1], ctr[2], ctr[3]}};
    U32x4 r = philox4x32_10(c, key[0], key[1]);
    for (int i = 0; i < 4; ++i) {
        out[i] = r.v[i];
    }
}

double qmb_philox_uniform_host(uint64_t seed, uint64_t index) { return philox_uniform(seed, index); }

// ------------------------------------------------------------------------ prepared objects ----
int qmb_terms_workspace_size(int n_qubytes, int64_t n_terms, size_t* bytes) {
    clear_error();
    QMB_TRY(check_qubytes(n_qubytes));
    if (!bytes) {
        return set_error(QMB_ERR_ARGUMENT, "bytes must be non-null");
    }
    *bytes = terms_bytes(n_qubytes, n_terms);
    return *bytes ? QMB_OK : QMB_ERR_ARGUMENT;
}

int qmb_terms_build(int n_qubytes, int particle_cut, const int16_t* site, const uint8_t* kind, const double* coef, int64_t n_terms,
                    void* buffer, size_t bytes, qmb_terms* out, void* stream) {
    clear_error();
    QMB_TRY(check_qubytes(n_qubytes));
    QMB_TRY(check_particle_cut(particle_cut));
    QMB_TRY(check_ptr(site, n_terms, "site"));
    QMB_TRY(check_ptr(kind, n_terms, "kind"));
    QMB_TRY(check_ptr(coef, n_terms, "coef"));
    if (!out || !buffer) {
        return set_error(QMB_ERR_ARGUMENT, "buffer and out must be non-null");
    }
    return terms_build(n_qubytes, particle_cut, site, kind, coef, n_terms, buffer, bytes, out, as_stream(stream));
}

int qmb_targets_workspace_size(int n_qubytes, int64_t n, size_t* bytes) {
    clear_error();
    QMB_TRY(check_qubytes(n_qubytes));
    if (!bytes) {
        return set_error(QMB_ERR_ARGUMENT, "bytes must be non-null");
    }
    *bytes = targets_bytes(n_qubytes, n);
    return *bytes ? QMB_OK : QMB_ERR_ARGUMENT;
}

int qmb_targets_build(int n_qubytes, const uint8_t* configs, int64_t n, void* buffer, size_t bytes, qmb_targets* out,
                      void* stream) {
    clear_error();
    QMB_TRY(check_ptr(configs, n, "configs"));
    if (!out || !buffer) {
        return set_error(QMB_ERR_ARGUMENT, "buffer and out must be non-null");
    }
    return targets_build(n_qubytes, configs, n, buffer, bytes, out, as_stream(stream));
}

int qmb_targets_build_keys(int n_qubytes, const uint8_t* configs, int64_t n, void* buffer, size_t bytes, qmb_targets* out,
                           void* stream) {
    clear_error();
    QMB_TRY(check_ptr(configs, n, "configs"));
    if (!out || !buffer) {
        return set_error(QMB_ERR_ARGUMENT, "buffer and out must be non-null");
    }
    return targets_build_keys(n_qubytes, configs, n, buffer, bytes, out, as_stream(stream));
}

int qmb_targets_build_index(const qmb_targets* targets, void* stream) {
    clear_error();
    if (!targets || !targets->buffer) {
        return set_error(QMB_ERR_ARGUMENT, "targets must be a descriptor from qmb_targets_build_keys");
    }
    return targets_build_index(targets, as_stream(stream));
}

int qmb_apply_prepared(const qmb_terms* terms, const qmb_targets* targets, const uint8_t* configs_i, int64_t n_i, const double* in,
                       double* out, int mode, const qmb_launch* launch, void* stream) {
    clear_error();
    QMB_TRY(check_ptr(configs_i, n_i, "configs_i"));
    return apply_prepared(terms, targets, configs_i, n_i, in, out, mode, launch, as_stream(stream));
}

// ---------------------------------------------------------------------------- pair path --
int qmb_pairs_workspace_size(int n_qubytes, size_t* bytes) {
    clear_error();
    if (!bytes) {
        return set_error(QMB_ERR_ARGUMENT, "bytes must be non-null");
    }
    QMB_TRY(check_qubytes(n_qubytes));
    *bytes = pairs_bytes(n_qubytes);
    return *bytes ? QMB_OK : QMB_ERR_UNSUPPORTED;
}

int qmb_pairs_build(int n_qubytes, int particle_cut, const int16_t* site, const uint8_t* kind, const double* coef, int64_t n_terms,
                    void* buffer, size_t bytes, qmb_pairs* out, void* stream) {
    clear_error();
    QMB_TRY(check_qubytes(n_qubytes));
    QMB_TRY(check_particle_cut(particle_cut));
    QMB_TRY(check_ptr(site, n_terms, "site"));
    QMB_TRY(check_ptr(kind, n_terms, "kind"));
    QMB_TRY(check_ptr(coef, n_terms, "coef"));
    if (!out || !buffer) {
        return set_error(QMB_ERR_ARGUMENT, "buffer and out must be non-null");
    }
    return pairs_build(n_qubytes, particle_cut, site, kind, coef, n_terms, buffer, bytes, out, as_stream(stream));
}

int qmb_pairs_set_real(qmb_pairs* pairs, int32_t real, void* stream) {
    clear_error();
    return pairs_set_real(pairs, real, as_stream(stream));
}

int qmb_singles_set_real(qmb_singles* singles, int32_t real, void* stream) {
    clear_error();
    return singles_set_real(singles, real, as_stream(stream));
}

int qmb_singles_workspace_size(int n_qubytes, size_t* bytes) {
    clear_error();
    if (!bytes) {
        return set_error(QMB_ERR_ARGUMENT, "bytes must be non-null");
    }
    QMB_TRY(check_qubytes(n_qubytes));
    *bytes = singles_bytes(n_qubytes);
    return *bytes ? QMB_OK : QMB_ERR_UNSUPPORTED;
}

int qmb_singles_build(int n_qubytes, int particle_cut, const int16_t* site, const uint8_t* kind, const double* coef,
                      int64_t n_terms, void* buffer, size_t bytes, qmb_singles* out, void* stream) {
    clear_error();
    QMB_TRY(check_qubytes(n_qubytes));
    QMB_TRY(check_particle_cut(particle_cut));
    QMB_TRY(check_ptr(site, n_terms, "site"));
    QMB_TRY(check_ptr(kind, n_terms, "kind"));
    QMB_TRY(check_ptr(coef, n_terms, "coef"));
    if (!out || !buffer) {
        return set_error(QMB_ERR_ARGUMENT, "buffer and out must be non-null");
    }
    return singles_build(n_qubytes, particle_cut, site, kind, coef, n_terms, buffer, bytes, out, as_stream(stream));
}

int qmb_terms_build_residual(int n_qubytes, int particle_cut, const int16_t* site, const uint8_t* kind, const double* coef,
                             int64_t n_terms, int32_t exclude, void* buffer, size_t bytes, qmb_terms* out, void* stream) {
    clear_error();
    QMB_TRY(check_qubytes(n_qubytes));
    QMB_TRY(check_particle_cut(particle_cut));
    QMB_TRY(check_ptr(site, n_terms, "site"));
    QMB_TRY(check_ptr(kind, n_terms, "kind"));
    QMB_TRY(check_ptr(coef, n_terms, "coef"));
    if (!out || !buffer) {
        return set_error(QMB_ERR_ARGUMENT, "buffer and out must be non-null");
    }
    if (exclude < 0 || exclude > (kExcludePairs | kExcludeSingles | kExcludeDiagonal | kExcludeHops)) {
        return set_error(QMB_ERR_ARGUMENT, "exclude must be a mask of 1, 2, 4 and 8, got %d", exclude);
    }
    return terms_build(n_qubytes, particle_cut, site, kind, coef, n_terms, buffer, bytes, out, as_stream(stream), exclude);
}

int qmb_cores_workspace_size(int n_qubytes, int64_t n, int32_t max_occupied, int32_t order, size_t* bytes) {
    clear_error();
    if (!bytes) {
        return set_error(QMB_ERR_ARGUMENT, "bytes must be non-null");
    }
    *bytes = cores_bytes(n_qubytes, n, max_occupied, order);
    return *bytes ? QMB_OK : QMB_ERR_ARGUMENT;
}

int qmb_cores_build(const qmb_targets* targets, int32_t max_occupied, int32_t order, void* buffer, size_t bytes, qmb_cores* out,
                    void* stream) {
    clear_error();
    if (!out || !buffer) {
        return set_error(QMB_ERR_ARGUMENT, "buffer and out must be non-null");
    }
    return cores_build(targets, max_occupied, order, buffer, bytes, out, as_stream(stream));
}

int qmb_apply_pairs(const qmb_pairs* pairs, const qmb_cores* cores, const qmb_targets* targets, const double* in, double* out,
                    int mode, void* stream) {
    clear_error();
    if (cores && cores->n > 0 && (!in || !out)) {
        return set_error(QMB_ERR_ARGUMENT, "in and out must be non-null");
    }
    return apply_pairs(pairs, cores, targets, in, out, mode, as_stream(stream));
}

int qmb_apply_singles(const qmb_singles* singles, const qmb_cores* cores, const qmb_targets* targets, const double* in,
                      double* out, int mode, void* stream) {
    clear_error();
    if (cores && cores->n > 0 && (!in || !out)) {
        return set_error(QMB_ERR_ARGUMENT, "in and out must be non-null");
    }
    return apply_singles(singles, cores, targets, in, out, mode, as_stream(stream));
}

int qmb_apply_diagonal(const qmb_pairs* pairs, const qmb_targets* targets, const double* in, double* out, int mode, void* stream) {
    clear_error();
    if (targets && targets->n > 0 && (!in || !out)) {
        return set_error(QMB_ERR_ARGUMENT, "in and out must be non-null");
    }
    return apply_diagonal(pairs, targets, in, out, mode, as_stream(stream));
}

int qmb_strings_workspace_size(int n_qubytes, int64_t n, size_t* bytes) {
    clear_error();
    if (!bytes) {
        return set_error(QMB_ERR_ARGUMENT, "bytes must be non-null");
    }
    *bytes = strings_bytes(n_qubytes, n);
    return *bytes ? QMB_OK : QMB_ERR_ARGUMENT;
}

int qmb_strings_build(const qmb_targets* targets, int32_t parity, void* buffer, size_t bytes, qmb_strings* out, void* stream) {
    clear_error();
    if (!buffer) {
        return set_error(QMB_ERR_ARGUMENT, "buffer must be non-null");
    }
    return strings_build(targets, parity, buffer, bytes, out, as_stream(stream));
}

int qmb_apply_hops(const qmb_singles* singles, const qmb_strings* strings, const qmb_targets* targets, const double* in, double* out,
                   int mode, void* stream) {
    clear_error();
    return apply_hops(singles, strings, targets, in, out, mode, as_stream(stream));
}

int qmb_apply_diagonal_terms(const qmb_terms* terms, const qmb_targets* targets, const double* in, double* out, int mode,
                             void* stream) {
    clear_error();
    return apply_diagonal_terms(terms, targets, in, out, mode, as_stream(stream));
}

int qmb_diagonal_values(const qmb_terms* terms, const qmb_targets* targets, double* d, void* stream) {
    clear_error();
    return diagonal_values(terms, targets, d, as_stream(stream));
}

int qmb_apply_diagonal_values(int64_t n, const double* d, const double* in, double* out, int mode, void* stream) {
    clear_error();
    return apply_diagonal_values(n, d, in, out, mode, as_stream(stream));
}

int qmb_apply_structured_dual(const qmb_pairs* pairs, const qmb_singles* singles, const qmb_cores* cores2, const qmb_cores* cores1,
                              const qmb_targets* targets, const double* in, double* out, const double* in2, double* out2, int mode,
                              void* stream) {
    clear_error();
    return apply_structured_dual(pairs, singles, cores2, cores1, targets, in, out, in2, out2, mode, as_stream(stream));
}

int qmb_batch_workspace_size(int n_items, size_t* bytes) {
    clear_error();
    if (!bytes) {
        return set_error(QMB_ERR_ARGUMENT, "bytes must be non-null");
    }
    *bytes = batch_bytes(n_items);
    return *bytes ? QMB_OK : QMB_ERR_ARGUMENT;
}

int qmb_batch_plan(const qmb_batch_item* items, int n_items, void* buffer, size_t bytes, qmb_batch* out, void* stream) {
    clear_error();
    if (!items || !out || !buffer) {
        return set_error(QMB_ERR_ARGUMENT, "items, buffer and out must be non-null");
    }
    return batch_plan(items, n_items, buffer, bytes, out, as_stream(stream));
}

int qmb_apply_batch(const qmb_batch* batch, int mode, int waves, void* stream) {
    clear_error();
    return apply_batch(batch, mode, waves, as_stream(stream));
}

int qmb_sampler_children(int dtype, int64_t n, int site, int n_sites, int spin_up, int spin_down, const uint64_t* prefix, int words,
                         const void* logits, const void* unperturbed, const void* perturbed, const void* uniform, uint64_t seed,
                         void* child_unperturbed, void* child_perturbed, void* stream) {
    clear_error();
    return sampler_children(dtype, n, site, n_sites, spin_up, spin_down, prefix, words, logits, unperturbed, perturbed, uniform, seed,
                            child_unperturbed, child_perturbed, as_stream(stream));
}

int qmb_sampler_select(int dtype, int64_t n_sel, const int64_t* index, int64_t n_parent, int site, const uint64_t* prefix, int words,
                       const void* child_unperturbed, const void* child_perturbed, uint64_t* out_prefix, void* out_unperturbed,
                       void* out_perturbed, void* stream) {
    clear_error();
    return sampler_select(dtype, n_sel, index, n_parent, site, prefix, words, child_unperturbed, child_perturbed, out_prefix,
                          out_unperturbed, out_perturbed, as_stream(stream));
}

int qmb_sampler_prefix_inputs(int dtype, int64_t n, int site, const uint64_t* prefix, int words, void* out, void* stream) {
    clear_error();
    return sampler_prefix_inputs(dtype, n, site, prefix, words, out, as_stream(stream));
}

int qmb_sampler_prefix_configs(int64_t n, int n_sites, const uint64_t* prefix, int words, const int64_t* ordering, int n_qubytes,
                               uint8_t* out, void* stream) {
    clear_error();
    return sampler_prefix_configs(n, n_sites, prefix, words, ordering, n_qubytes, out, as_stream(stream));
}

int qmb_network_inputs(int dtype, int64_t n, int n_sites, const uint8_t* configs, int n_qubytes, const int64_t* ordering_reversed,
                       void* out, void* stream) {
    clear_error();
    return network_inputs(dtype, n, n_sites, configs, n_qubytes, ordering_reversed, out, as_stream(stream));
}

int qmb_network_amplitude(int dtype, int64_t n, int n_sites, int spin_up, int spin_down, const uint8_t* configs, int n_qubytes,
                          const int64_t* ordering_reversed, const void* logits, double* out, const double* grad, void* grad_logits,
                          void* stream) {
    clear_error();
    return network_amplitude(dtype, n, n_sites, spin_up, spin_down, configs, n_qubytes, ordering_reversed, logits, out, grad, grad_logits,
                             as_stream(stream));
}

int qmb_network_fused_supported(int n_sites, int hidden) { return network_fused_supported(n_sites, hidden); }

int qmb_network_forward_fused(int64_t n, int n_sites, int hidden, int spin_up, int spin_down, const uint8_t* configs, int n_qubytes,
                              const int64_t* ordering_reversed, const float* w1, const float* b1, const float* w2, const float* b2,
                              const float* wp1, const float* bp1, const float* wp2, const float* bp2, double* amplitude, float* phase,
                              void* stream) {
    clear_error();
    return network_forward_fused(n, n_sites, hidden, spin_up, spin_down, configs, n_qubytes, ordering_reversed, w1, b1, w2, b2, wp1, bp1, wp2,
                                 bp2, amplitude, phase, as_stream(stream));
}

int qmb_sampler_step_fused(int64_t n, int site, int n_sites, int hidden, int spin_up, int spin_down, const uint64_t* prefix, int words,
                           const float* w1, const float* b1, const float* w2, const float* b2, const float* unperturbed,
                           const float* perturbed, const float* uniform, uint64_t seed, float* child_unperturbed, float* child_perturbed,
                           void* stream) {
    clear_error();
    return sampler_step_fused(n, site, n_sites, hidden, spin_up, spin_down, prefix, words, w1, b1, w2, b2, unperturbed, perturbed, uniform,
                              seed, child_unperturbed, child_perturbed, as_stream(stream));
}

int qmb_targets_lookup(const qmb_targets* targets, const uint8_t* configs, int64_t n, int64_t* rows, void* stream) {
    clear_error();
    return targets_lookup(targets, configs, n, rows, as_stream(stream));
}

int qmb_lanczos_workspace_size(size_t* bytes) {
    clear_error();
    if (!bytes) {
        return set_error(QMB_ERR_ARGUMENT, "bytes must be non-null");
    }
    *bytes = lanczos_bytes();
    return QMB_OK;
}

int qmb_lanczos_step(int64_t n, const double* v, const double* v_prev, double beta, double* w, double* v_next, double* scalars,
                     void* workspace, size_t bytes, void* stream) {
    clear_error();
    return lanczos_step(n, v, v_prev, beta, w, v_next, scalars, workspace, bytes, as_stream(stream));
}

int qmb_single_excitation_host(uint64_t rec, int particle_cut, int32_t out[5]) {
    clear_error();
    SingleExcitation e;
    if (!single_excitation(rec, particle_cut, &e)) {
        return 0;
    }
    if (out) {
        out[0] = e.r;
        out[1] = e.p;
        out[2] = e.k;
        out[3] = e.kocc;
        out[4] = e.delta;
    }
    return 1;
}

int qmb_pair_excitation_host(uint64_t rec, int particle_cut, int32_t out[5]) {
    clear_error();
    PairExcitation e;
    if (!pair_excitation(rec, particle_cut, &e)) {
        return 0;
    }
    if (out) {
        out[0] = e.r0;
        out[1] = e.r1;
        out[2] = e.c0;
        out[3] = e.c1;
        out[4] = e.delta;
    }
    return 1;
}

// ------------------------------------------------------------------------------ apply_within --
int qmb_apply_within_workspace_size(int n_qubytes, int64_t n_i, int64_t n_j, int64_t n_terms, size_t* bytes) {
    clear_error();
    (void)n_i;
    QMB_TRY(check_qubytes(n_qubytes));
    if (!bytes) {
        return set_error(QMB_ERR_ARGUMENT, "bytes must be non-null");
    }
    *bytes = terms_bytes(n_qubytes, n_terms) + targets_bytes(n_qubytes, n_j) + 512;
    return QMB_OK;
}

}  // extern "C"

// ------------------------------------------------------------- structured path behind the C ABI --
// qmb_apply_within / qmb_apply_within_adjoint on ONE config set (configs_i and configs_j the same
// buffer) take the structured path -- pair and single excitations applied by walking core groups,
// diagonal terms from the folded table, the rest enumerated -- when the table has pair excitations,
// the rows are distinct and the workspace holds the core indexes (qmb_apply_within_structured_
// workspace_size).  Choosing needs four numbers from the device (pair terms, rows repeated, the
// most occupied sites in a row, coefficients with an imaginary part); they are read back with
// one synchronisation of the stream, as the reference's op synchronises after every call.
namespace qmb {
namespace {

struct StructuredLayout {
    size_t terms, targets, pairs, singles, residual, cores2, cores1, total;
};

StructuredLayout structured_layout(int nq, int64_t n, int64_t n_terms, int max_occupied) {
    StructuredLayout l{};
    l.terms = terms_bytes(nq, n_terms);
    l.targets = targets_bytes(nq, n);
    l.pairs = pairs_bytes(nq);
    l.singles = singles_bytes(nq);
    l.residual = terms_bytes(nq, n_terms);
    l.cores2 = cores_bytes(nq, n, max_occupied, 2);
    l.cores1 = cores_bytes(nq, n, max_occupied, 1);
    l.total = l.terms + l.targets + l.pairs + l.singles + l.residual + l.cores2 + l.cores1 + 8 * 256 + 1024;
    return l;
}

// plan[0] pair terms, [1] single terms, [2] repeated rows, [3] most occupied sites in a row,
// [4] coefficients with a nonzero imaginary part, [5] terms outside the three structured classes
__global__ void plan_zero_kernel(int64_t* plan) {
    if (threadIdx.x < 8) {
        plan[threadIdx.x] = 0;
    }
}

__global__ void plan_scan_kernel(const uint8_t* __restrict__ configs, int64_t n, int nq, const double* __restrict__ coef,
                                 int64_t n_terms, int64_t* __restrict__ plan) {
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    int occ = 0;
    if (i < n) {
        const uint8_t* row = configs + i * nq;
        for (int b = 0; b < nq; ++b) {
            occ += __popc(static_cast<unsigned>(row[b]));
        }
    }
    // one atomic per wave for the maximum
    for (int off = 32; off > 0; off >>= 1) {
        occ = max(occ, __shfl_xor(occ, off));
    }
    if ((threadIdx.x & 63) == 0 && occ > 0) {
        atomicMax(reinterpret_cast<unsigned long long*>(&plan[3]), static_cast<unsigned long long>(occ));
    }
    const unsigned long long imag = (i < n_terms && coef[2 * i + 1] != 0.0) ? 1ull : 0ull;
    const unsigned long long waves = __popcll(__ballot(imag != 0ull));
    if ((threadIdx.x & 63) == 0 && waves > 0) {
        atomicAdd(reinterpret_cast<unsigned long long*>(&plan[4]), waves);
    }
}

__global__ void plan_gather_kernel(const int64_t* __restrict__ pair_counts, const int64_t* __restrict__ single_counts,
                                   const int64_t* __restrict__ duplicates, const int64_t* __restrict__ residual_counts,
                                   int64_t* __restrict__ plan) {
    if (threadIdx.x == 0) {
        plan[0] = pair_counts[0];
        plan[1] = single_counts[0];
        plan[2] = duplicates[0];
        plan[5] = residual_counts[0];
    }
}

// 0: applied through the structured path; 1: not applicable (the caller enumerates); else an error.
int apply_structured(int nq, int pc, const uint8_t* configs, int64_t n, const int16_t* site, const uint8_t* kind,
                     const double* coef, int64_t n_terms, const double* in, double* out, int mode, void* workspace,
                     size_t workspace_bytes, hipStream_t s) {
    // the smallest layout (two occupied sites) decides whether to look at all
    if (nq > kPairMaxQubytes || n <= 0 || n_terms <= 0 || workspace_bytes < structured_layout(nq, n, n_terms, 2).total) {
        return 1;
    }
    const StructuredLayout lo = structured_layout(nq, n, n_terms, 2);
    Regions r{static_cast<char*>(workspace), 0, workspace_bytes};
    void* terms_buf = r.next(lo.terms);
    qmb_targets targets{};
    QMB_TRY(targets_build(nq, configs, n, r.next(lo.targets), lo.targets, &targets, s));
    qmb_pairs pairs{};
    QMB_TRY(pairs_build(nq, pc, site, kind, coef, n_terms, r.next(lo.pairs), lo.pairs, &pairs, s));
    qmb_singles singles{};
    QMB_TRY(singles_build(nq, pc, site, kind, coef, n_terms, r.next(lo.singles), lo.singles, &singles, s));
    // the terms outside the three classes (none for a molecular table); excluding single
    // excitations is a no-op when the table has none
    qmb_terms residual{};
    QMB_TRY(terms_build(nq, pc, site, kind, coef, n_terms, r.next(lo.residual), lo.residual, &residual, s,
                        kExcludePairs | kExcludeDiagonal | kExcludeSingles));
    int64_t* plan = static_cast<int64_t*>(r.next(8 * sizeof(int64_t)));
    const size_t cores_off = r.off;
    plan_zero_kernel<<<1, 64, 0, s>>>(plan);
    QMB_CHECK_LAUNCH();
    const int64_t reach = std::max(n, n_terms);
    plan_scan_kernel<<<static_cast<unsigned>((reach + 255) / 256), 256, 0, s>>>(configs, n, nq, coef, n_terms, plan);
    QMB_CHECK_LAUNCH();
    plan_gather_kernel<<<1, 64, 0, s>>>(pairs.counts, singles.counts, targets.duplicates, residual.counts, plan);
    QMB_CHECK_LAUNCH();
    int64_t host[8] = {};
    QMB_HIP_TRY(hipMemcpyAsync(host, plan, 6 * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    QMB_HIP_TRY(hipStreamSynchronize(s));
    const int64_t n_pair = host[0], n_single = host[1], repeated = host[2], occ = host[3], n_imag = host[4], n_residual = host[5];
    if (n_pair == 0 || repeated != 0 || occ < 2) {
        return 1;
    }
    const StructuredLayout full = structured_layout(nq, n, n_terms, static_cast<int>(occ));
    if (workspace_bytes < full.total) {
        return 1;
    }
    (void)terms_buf;
    if (n_imag == 0) {
        QMB_TRY(pairs_set_real(&pairs, 1, s));
        if (n_single > 0) {
            QMB_TRY(singles_set_real(&singles, 1, s));
        }
    }
    r.off = cores_off;
    qmb_cores cores2{};
    QMB_TRY(cores_build(&targets, static_cast<int>(occ), 2, r.next(full.cores2), full.cores2, &cores2, s));
    qmb_cores cores1{};
    if (n_single > 0) {
        QMB_TRY(cores_build(&targets, static_cast<int>(occ), 1, r.next(full.cores1), full.cores1, &cores1, s));
    }
    QMB_TRY(launch_zero_f64(out, 2 * n, s));
    QMB_TRY(apply_diagonal(&pairs, &targets, in, out, mode, s));
    if (n_single > 0) {
        QMB_TRY(apply_singles(&singles, &cores1, &targets, in, out, mode, s));
    }
    QMB_TRY(apply_pairs(&pairs, &cores2, &targets, in, out, mode, s));
    if (n_residual > 0) {
        QMB_TRY(apply_prepared(&residual, &targets, configs, n, in, out, mode, nullptr, s));
    }
    return 0;
}

// The hop path behind the C ABI (one config set, same-parity hops): the full term table (its
// diagonal records), the target index (its duplicate count), the single-excitation table, the terms
// left to enumeration, and a strings index per parity.  Two synchronisations: one reads the term
// classes and repeated rows, one the strings indexes' pair-test counts, which must stay within
// kHopTestsPerTermRow per (row, hop term) or the call enumerates instead.
constexpr int64_t kHopTestsPerTermRow = 4;

struct HopLayout {
    size_t terms, targets, singles, residual, strings, total;
};

HopLayout hop_layout(int nq, int64_t n, int64_t n_terms) {
    HopLayout l{};
    l.terms = terms_bytes(nq, n_terms);
    l.targets = targets_bytes(nq, n);
    l.singles = singles_bytes(nq);
    l.residual = terms_bytes(nq, n_terms);
    l.strings = strings_bytes(nq, n);
    l.total = l.terms + l.targets + l.singles + l.residual + 2 * l.strings + 8 * 256 + 1024;
    return l;
}

// plan[0] valid terms, [1] diagonal terms, [2] terms left to enumeration, [3] repeated rows,
// [4] coefficients with a nonzero imaginary part (from plan_scan_kernel)
__global__ void hop_plan_gather_kernel(const int64_t* __restrict__ term_counts, const int64_t* __restrict__ residual_counts,
                                       const int64_t* __restrict__ duplicates, int64_t* __restrict__ plan) {
    if (threadIdx.x == 0) {
        plan[0] = term_counts[0];
        plan[1] = term_counts[1];
        plan[2] = residual_counts[0];
        plan[3] = duplicates[0];
    }
}

__global__ void hop_tests_kernel(const int64_t* __restrict__ counts0, const int64_t* __restrict__ counts1, int64_t* __restrict__ plan) {
    if (threadIdx.x == 0) {
        plan[5] = counts0[2] + counts1[2];
    }
}

// 0: applied through the hop path; 1: not applicable (the caller enumerates); else an error.
int apply_hop_path(int nq, int pc, const uint8_t* configs, int64_t n, const int16_t* site, const uint8_t* kind, const double* coef,
                   int64_t n_terms, const double* in, double* out, int mode, void* workspace, size_t workspace_bytes, hipStream_t s) {
    if (nq > kPairMaxQubytes || n <= 0 || n_terms <= 0 || n >= (int64_t{1} << 31)) {
        return 1;
    }
    const HopLayout lo = hop_layout(nq, n, n_terms);
    if (workspace_bytes < lo.total) {
        return 1;
    }
    Regions r{static_cast<char*>(workspace), 0, workspace_bytes};
    qmb_terms terms{};
    QMB_TRY(terms_build(nq, pc, site, kind, coef, n_terms, r.next(lo.terms), lo.terms, &terms, s));
    qmb_targets targets{};
    QMB_TRY(targets_build(nq, configs, n, r.next(lo.targets), lo.targets, &targets, s));
    qmb_singles singles{};
    QMB_TRY(singles_build(nq, pc, site, kind, coef, n_terms, r.next(lo.singles), lo.singles, &singles, s));
    qmb_terms residual{};
    QMB_TRY(terms_build(nq, pc, site, kind, coef, n_terms, r.next(lo.residual), lo.residual, &residual, s,
                        kExcludeDiagonal | kExcludeHops));
    int64_t* plan = static_cast<int64_t*>(r.next(8 * sizeof(int64_t)));
    plan_zero_kernel<<<1, 64, 0, s>>>(plan);
    QMB_CHECK_LAUNCH();
    // plan[3] and plan[4] of plan_scan (most occupied sites, imaginary coefficients); [3] is then
    // overwritten with the repeated rows
    const int64_t reach = std::max(n, n_terms);
    plan_scan_kernel<<<static_cast<unsigned>((reach + 255) / 256), 256, 0, s>>>(configs, n, nq, coef, n_terms, plan);
    QMB_CHECK_LAUNCH();
    hop_plan_gather_kernel<<<1, 64, 0, s>>>(terms.counts, residual.counts, targets.duplicates, plan);
    QMB_CHECK_LAUNCH();
    int64_t host[8] = {};
    QMB_HIP_TRY(hipMemcpyAsync(host, plan, 5 * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    QMB_HIP_TRY(hipStreamSynchronize(s));
    const int64_t n_hops = host[0] - host[1] - host[2];
    if (n_hops <= 0 || host[3] != 0) {
        return 1;
    }
    if (host[4] == 0) {
        QMB_TRY(singles_set_real(&singles, 1, s));
    }
    qmb_strings strings[2]{};
    for (int parity = 0; parity < 2; ++parity) {
        QMB_TRY(strings_build(&targets, parity, r.next(lo.strings), lo.strings, &strings[parity], s));
    }
    hop_tests_kernel<<<1, 64, 0, s>>>(strings[0].counts, strings[1].counts, plan);
    QMB_CHECK_LAUNCH();
    QMB_HIP_TRY(hipMemcpyAsync(host + 5, plan + 5, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    QMB_HIP_TRY(hipStreamSynchronize(s));
    if (host[5] > kHopTestsPerTermRow * n * n_hops) {
        return 1;
    }
    QMB_TRY(launch_zero_f64(out, 2 * n, s));
    QMB_TRY(apply_diagonal_terms(&terms, &targets, in, out, mode, s));
    for (int parity = 0; parity < 2; ++parity) {
        QMB_TRY(apply_hops(&singles, &strings[parity], &targets, in, out, mode, s));
    }
    if (host[2] > 0) {
        QMB_TRY(apply_prepared(&residual, &targets, configs, n, in, out, mode, nullptr, s));
    }
    return 0;
}

}  // namespace
}  // namespace qmb

extern "C" {

int qmb_apply_within_structured_workspace_size(int n_qubytes, int64_t n, int64_t n_terms, int32_t max_occupied, size_t* bytes) {
    clear_error();
    QMB_TRY(check_qubytes(n_qubytes));
    if (!bytes) {
        return set_error(QMB_ERR_ARGUMENT, "bytes must be non-null");
    }
    if (n < 0 || n_terms < 0 || max_occupied < 0 || max_occupied > 8 * n_qubytes) {
        return set_error(QMB_ERR_ARGUMENT, "n, n_terms must be non-negative and max_occupied in [0, %d]", 8 * n_qubytes);
    }
    if (n_qubytes > kPairMaxQubytes) {
        return set_error(QMB_ERR_UNSUPPORTED, "the structured path covers configs of up to %d bytes", kPairMaxQubytes);
    }
    *bytes = std::max(structured_layout(n_qubytes, n, n_terms, std::max(2, static_cast<int>(max_occupied))).total,
                      terms_bytes(n_qubytes, n_terms) + targets_bytes(n_qubytes, n) + 512);
    return QMB_OK;
}

int qmb_apply_within_hops_workspace_size(int n_qubytes, int64_t n, int64_t n_terms, size_t* bytes) {
    clear_error();
    QMB_TRY(check_qubytes(n_qubytes));
    if (!bytes) {
        return set_error(QMB_ERR_ARGUMENT, "bytes must be non-null");
    }
    if (n < 0 || n_terms < 0) {
        return set_error(QMB_ERR_ARGUMENT, "n and n_terms must be non-negative");
    }
    if (n_qubytes > kPairMaxQubytes) {
        return set_error(QMB_ERR_UNSUPPORTED, "the hop path covers configs of up to %d bytes", kPairMaxQubytes);
    }
    *bytes = std::max(hop_layout(n_qubytes, n, n_terms).total, terms_bytes(n_qubytes, n_terms) + targets_bytes(n_qubytes, n) + 512);
    return QMB_OK;
}

static int apply_oneshot(int n_qubytes, int particle_cut, const uint8_t* configs_i, int64_t n_i, const uint8_t* configs_j,
                         int64_t n_j, const int16_t* site, const uint8_t* kind, const double* coef, int64_t n_terms,
                         const double* in, double* out, int mode, void* workspace, size_t workspace_bytes, void* stream) {
    QMB_TRY(check_qubytes(n_qubytes));
    QMB_TRY(check_particle_cut(particle_cut));
    QMB_TRY(check_ptr(configs_i, n_i, "configs_i"));
    QMB_TRY(check_ptr(configs_j, n_j, "configs_j"));
    QMB_TRY(check_ptr(site, n_terms, "site"));
    QMB_TRY(check_ptr(kind, n_terms, "kind"));
    QMB_TRY(check_ptr(coef, n_terms, "coef"));
    const size_t tb = terms_bytes(n_qubytes, n_terms);
    const size_t gb = targets_bytes(n_qubytes, n_j);
    QMB_TRY(need_workspace(tb + gb + 512, workspace_bytes, workspace));
    hipStream_t s = as_stream(stream);
    if (configs_i == configs_j && n_i == n_j) {
        int rc = apply_structured(n_qubytes, particle_cut, configs_i, n_i, site, kind, coef, n_terms, in, out, mode, workspace,
                                  workspace_bytes, s);
        if (rc != 1) {
            return rc;
        }
        rc = apply_hop_path(n_qubytes, particle_cut, configs_i, n_i, site, kind, coef, n_terms, in, out, mode, workspace,
                            workspace_bytes, s);
        if (rc != 1) {
            return rc;
        }
    }
    const int64_t n_out = mode == 0 ? n_j : n_i;
    QMB_TRY(launch_zero_f64(out, 2 * n_out, s));
    if (n_i == 0 || n_j == 0 || n_terms == 0) {
        return QMB_OK;
    }
    Regions r{static_cast<char*>(workspace), 0, workspace_bytes};
    qmb_terms terms{};
    qmb_targets targets{};
    QMB_TRY(terms_build(n_qubytes, particle_cut, site, kind, coef, n_terms, r.next(tb), tb, &terms, s));
    QMB_TRY(targets_build(n_qubytes, configs_j, n_j, r.next(gb), gb, &targets, s));
    return apply_prepared(&terms, &targets, configs_i, n_i, in, out, mode, nullptr, s);
}

int qmb_apply_within(int n_qubytes, int particle_cut, const uint8_t* configs_i, const double* psi_i, int64_t n_i,
                     const uint8_t* configs_j, int64_t n_j, const int16_t* site, const uint8_t* kind, const double* coef,
                     int64_t n_terms, double* psi_j, void* workspace, size_t workspace_bytes, void* stream) {
    clear_error();
    QMB_TRY(check_ptr(psi_i, n_i, "psi_i"));
    QMB_TRY(check_ptr(psi_j, n_j, "psi_j"));
    return apply_oneshot(n_qubytes, particle_cut, configs_i, n_i, configs_j, n_j, site, kind, coef, n_terms, psi_i, psi_j, 0,
                         workspace, workspace_bytes, stream);
}

int qmb_apply_within_adjoint(int n_qubytes, int particle_cut, const uint8_t* configs_i, int64_t n_i, const uint8_t* configs_j,
                             const double* grad_j, int64_t n_j, const int16_t* site, const uint8_t* kind, const double* coef,
                             int64_t n_terms, double* grad_i, void* workspace, size_t workspace_bytes, void* stream) {
    clear_error();
    QMB_TRY(check_ptr(grad_j, n_j, "grad_j"));
    QMB_TRY(check_ptr(grad_i, n_i, "grad_i"));
    return apply_oneshot(n_qubytes, particle_cut, configs_i, n_i, configs_j, n_j, site, kind, coef, n_terms, grad_j, grad_i, 1,
                         workspace, workspace_bytes, stream);
}

// --------------------------------------------------------------------- no-allocation variant --
// qmb_apply_within for callers that cannot allocate: the term table and the target index go into a
// file-scope device array instead of a caller workspace.  Each device holds its own copy of the
// array for the life of the process; its address on a device is looked up once, with that device
// current, and only that address is handed to kernels and copies.
}  // extern "C"

namespace qmb {
namespace {
constexpr size_t kStaticScratchBytes = size_t{64} << 20;
constexpr int kMaxDevices = 64;
__device__ __align__(256) unsigned char scratch[kStaticScratchBytes];
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

int static_scratch(void** out) {
    int dev = 0;
    QMB_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        return set_error(QMB_ERR_UNSUPPORTED, "device %d is beyond the %d devices the static scratch serves", dev, kMaxDevices);
    }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[dev] == nullptr) {
        void* ptr = nullptr;
        QMB_HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (ptr == nullptr) {
            return set_error(QMB_ERR_HIP, "hipGetSymbolAddress returned no address for the static scratch");
        }
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return QMB_OK;
}
}  // namespace
}  // namespace qmb

extern "C" {

size_t qmb_static_scratch_bytes(void) { return kStaticScratchBytes; }

int qmb_apply_within_noalloc(int n_qubytes, int particle_cut, const uint8_t* configs_i, const double* psi_i, int64_t n_i,
                             const uint8_t* configs_j, int64_t n_j, const int16_t* site, const uint8_t* kind, const double* coef,
                             int64_t n_terms, double* psi_j, void* stream) {
    clear_error();
    QMB_TRY(check_qubytes(n_qubytes));
    QMB_TRY(check_particle_cut(particle_cut));
    QMB_TRY(check_ptr(psi_i, n_i, "psi_i"));
    QMB_TRY(check_ptr(psi_j, n_j, "psi_j"));
    QMB_TRY(check_ptr(configs_i, n_i, "configs_i"));
    QMB_TRY(check_ptr(configs_j, n_j, "configs_j"));
    if (n_terms < 0) {
        return set_error(QMB_ERR_ARGUMENT, "n_terms must be non-negative");
    }
    const size_t need = terms_bytes(n_qubytes, n_terms) + targets_bytes(n_qubytes, n_j) + 512;
    if (need > kStaticScratchBytes) {
        return set_error(QMB_ERR_WORKSPACE,
                         "qmb_apply_within_noalloc: %lld terms and %lld target configs need %zu bytes of scratch, above its "
                         "fixed %zu bytes; use qmb_apply_within with a workspace",
                         static_cast<long long>(n_terms), static_cast<long long>(n_j), need, kStaticScratchBytes);
    }
    void* ptr = nullptr;
    QMB_TRY(static_scratch(&ptr));
    return apply_oneshot(n_qubytes, particle_cut, configs_i, n_i, configs_j, n_j, site, kind, coef, n_terms, psi_i, psi_j, 0, ptr,
                         kStaticScratchBytes, stream);
}

// ------------------------------------------------------------------------------------ plain --
int qmb_apply_within_plain_workspace_size(int n_qubytes, int64_t n_j, size_t* bytes) {
    clear_error();
    QMB_TRY(check_qubytes(n_qubytes));
    if (!bytes) {
        return set_error(QMB_ERR_ARGUMENT, "bytes must be non-null");
    }
    const int W = n_words_of(n_qubytes);
    const size_t N = static_cast<size_t>(n_j > 0 ? n_j : 1);
    *bytes = N * W * 8 + 256 + N * 4 + 256 + sort_configs_scratch_bytes(W, n_j) + 512;
    return QMB_OK;
}

int qmb_apply_within_plain(int n_qubytes, int particle_cut, const uint8_t* configs_i, const double* psi_i, int64_t n_i,
                           const uint8_t* configs_j, int64_t n_j, const int16_t* site, const uint8_t* kind, const double* coef,
                           int64_t n_terms, double* psi_j, void* workspace, size_t workspace_bytes, void* stream) {
    clear_error();
    QMB_TRY(check_qubytes(n_qubytes));
    QMB_TRY(check_particle_cut(particle_cut));
    QMB_TRY(check_ptr(configs_i, n_i, "configs_i"));
    QMB_TRY(check_ptr(psi_i, n_i, "psi_i"));
    QMB_TRY(check_ptr(configs_j, n_j, "configs_j"));
    QMB_TRY(check_ptr(psi_j, n_j, "psi_j"));
    QMB_TRY(check_ptr(site, n_terms, "site"));
    size_t need = 0;
    QMB_TRY(qmb_apply_within_plain_workspace_size(n_qubytes, n_j, &need));
    QMB_TRY(need_workspace(need, workspace_bytes, workspace));
    // The plain kernel reads raw (site, kind) ops, so reject sites it would mis-handle.
    hipStream_t s = as_stream(stream);
    QMB_TRY(launch_zero_f64(psi_j, 2 * n_j, s));
    if (n_i == 0 || n_j == 0 || n_terms == 0) {
        return QMB_OK;
    }
    const int W = n_words_of(n_qubytes);
    const size_t N = static_cast<size_t>(n_j);
    Regions r{static_cast<char*>(workspace), 0, workspace_bytes};
    uint64_t* keys_be = static_cast<uint64_t*>(r.next(N * W * 8));
    uint32_t* perm = static_cast<uint32_t*>(r.next(N * 4));
    const size_t sb = sort_configs_scratch_bytes(W, n_j);
    void* scratch = r.next(sb);
    QMB_TRY(sort_configs(n_qubytes, configs_j, n_j, keys_be, perm, scratch, sb, s));
    return apply_plain(n_qubytes, particle_cut, configs_i, psi_i, n_i, keys_be, perm, n_j, site, kind, coef, n_terms, psi_j, s);
}

// --------------------------------------------------------------------------- diagonal_term ----
int qmb_diagonal_term_workspace_size(int n_qubytes, int64_t n, int64_t n_terms, size_t* bytes) {
    clear_error();
    (void)n;
    QMB_TRY(check_qubytes(n_qubytes));
    if (!bytes) {
        return set_error(QMB_ERR_ARGUMENT, "bytes must be non-null");
    }
    *bytes = terms_bytes(n_qubytes, n_terms) + 256;
    return QMB_OK;
}

int qmb_diagonal_term(int n_qubytes, int particle_cut, const uint8_t* configs, int64_t n, const int16_t* site, const uint8_t* kind,
                      const double* coef, int64_t n_terms, double* out, void* workspace, size_t workspace_bytes, void* stream) {
    clear_error();
    QMB_TRY(check_qubytes(n_qubytes));
    QMB_TRY(check_particle_cut(particle_cut));
    QMB_TRY(check_ptr(configs, n, "configs"));
    QMB_TRY(check_ptr(out, n, "out"));
    QMB_TRY(check_ptr(site, n_terms, "site"));
    QMB_TRY(check_ptr(kind, n_terms, "kind"));
    QMB_TRY(check_ptr(coef, n_terms, "coef"));
    const size_t tb = terms_bytes(n_qubytes, n_terms);
    QMB_TRY(need_workspace(tb, workspace_bytes, workspace));
    hipStream_t s = as_stream(stream);
    Regions r{static_cast<char*>(workspace), 0, workspace_bytes};
    qmb_terms terms{};
    QMB_TRY(terms_build(n_qubytes, particle_cut, site, kind, coef, n_terms, r.next(tb), tb, &terms, s));
    return diagonal_prepared(&terms, configs, n, out, s);
}

// ------------------------------------------------------------------------- single_relative ----
int qmb_single_relative_workspace_size(int n_qubytes, int64_t n, int64_t n_terms, size_t* bytes) {
    clear_error();
    QMB_TRY(check_qubytes(n_qubytes));
    if (!bytes) {
        return set_error(QMB_ERR_ARGUMENT, "bytes must be non-null");
    }
    *bytes = targets_bytes(n_qubytes, n) + single_relative_scratch_bytes(n_terms) + 512;
    return QMB_OK;
}

int qmb_single_relative(int n_qubytes, int particle_cut, const uint8_t* configs, int64_t n, const int16_t* site, const uint8_t* kind,
                        const double* coef, int64_t n_terms, uint64_t seed, const double* keys, uint8_t* out, void* workspace,
                        size_t workspace_bytes, void* stream) {
    clear_error();
    QMB_TRY(check_qubytes(n_qubytes));
    QMB_TRY(check_particle_cut(particle_cut));
    QMB_TRY(check_ptr(configs, n, "configs"));
    QMB_TRY(check_ptr(out, n, "out"));
    QMB_TRY(check_ptr(site, n_terms, "site"));
    QMB_TRY(check_ptr(kind, n_terms, "kind"));
    QMB_TRY(check_ptr(coef, n_terms, "coef"));
    const size_t gb = targets_bytes(n_qubytes, n);
    const size_t sb = single_relative_scratch_bytes(n_terms);
    QMB_TRY(need_workspace(gb + sb + 512, workspace_bytes, workspace));
    hipStream_t s = as_stream(stream);
    Regions r{static_cast<char*>(workspace), 0, workspace_bytes};
    qmb_targets exclude{};
    QMB_TRY(targets_build(n_qubytes, configs, n, r.next(gb), gb, &exclude, s));
    return single_relative_prepared(n_qubytes, particle_cut, configs, n, site, kind, coef, n_terms, seed, keys, &exclude, out,
                                    r.next(sb), sb, s);
}

// --------------------------------------------------------------------------- find_relative ----
int qmb_find_relative_workspace_size(int n_qubytes, int64_t n_i, int64_t n_exclude, int64_t n_terms, int64_t count_selected,
                                     size_t* bytes) {
    clear_error();
    QMB_TRY(check_qubytes(n_qubytes));
    if (!bytes) {
        return set_error(QMB_ERR_ARGUMENT, "bytes must be non-null");
    }
    if (count_selected <= 0) {
        return set_error(QMB_ERR_ARGUMENT, "count_selected must be positive");
    }
    *bytes = terms_bytes(n_qubytes, n_terms) + targets_bytes(n_qubytes, n_exclude) +
             find_relative_scratch_bytes(n_qubytes, n_i, count_selected, n_terms) + 768;
    return QMB_OK;
}

int qmb_find_relative(int n_qubytes, int particle_cut, const uint8_t* configs_i, const double* psi_i, int64_t n_i,
                      int64_t count_selected, const int16_t* site, const uint8_t* kind, const double* coef, int64_t n_terms,
                      const uint8_t* configs_exclude, int64_t n_exclude, uint8_t* out, int64_t* out_rows, void* workspace,
                      size_t workspace_bytes, void* stream) {
    clear_error();
    QMB_TRY(check_qubytes(n_qubytes));
    QMB_TRY(check_particle_cut(particle_cut));
    QMB_TRY(check_ptr(configs_i, n_i, "configs_i"));
    QMB_TRY(check_ptr(psi_i, n_i, "psi_i"));
    QMB_TRY(check_ptr(site, n_terms, "site"));
    QMB_TRY(check_ptr(kind, n_terms, "kind"));
    QMB_TRY(check_ptr(coef, n_terms, "coef"));
    QMB_TRY(check_ptr(configs_exclude, n_exclude, "configs_exclude"));
    if (!out || !out_rows) {
        return set_error(QMB_ERR_ARGUMENT, "out and out_rows must be non-null");
    }
    size_t need = 0;
    QMB_TRY(qmb_find_relative_workspace_size(n_qubytes, n_i, n_exclude, n_terms, count_selected, &need));
    QMB_TRY(need_workspace(need, workspace_bytes, workspace));
    hipStream_t s = as_stream(stream);
    const size_t tb = terms_bytes(n_qubytes, n_terms);
    const size_t gb = targets_bytes(n_qubytes, n_exclude);
    const size_t fb = find_relative_scratch_bytes(n_qubytes, n_i, count_selected, n_terms);
    Regions r{static_cast<char*>(workspace), 0, workspace_bytes};
    qmb_terms terms{};
    qmb_targets exclude{};
    QMB_TRY(terms_build(n_qubytes, particle_cut, site, kind, coef, n_terms, r.next(tb), tb, &terms, s));
    QMB_TRY(targets_build(n_qubytes, configs_exclude, n_exclude, r.next(gb), gb, &exclude, s));
    return find_relative_prepared(&terms, &exclude, configs_i, psi_i, n_i, count_selected, out, out_rows, r.next(fb), fb, s);
}

}  // extern "C"
