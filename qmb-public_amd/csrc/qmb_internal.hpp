// qmb_internal.hpp -- declarations shared by the qmb_hip translation units (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstddef>
#include <cstdint>
#include <string>
#include <type_traits>

#include "qmb_common.hpp"
#include "qmb_hip.h"

namespace qmb {

// ---------------------------------------------------------------------------------------------
// Errors: the C ABI returns a code and keeps the message per host thread.
// ---------------------------------------------------------------------------------------------
int set_error(int code, const char* fmt, ...);
void clear_error();

#define QMB_HIP_TRY(expr)                                                                                  \
    do {                                                                                                   \
        hipError_t qmb_e_ = (expr);                                                                        \
        if (qmb_e_ != hipSuccess) {                                                                        \
            return ::qmb::set_error(QMB_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(qmb_e_), \
                                    __FILE__, __LINE__);                                                   \
        }                                                                                                  \
    } while (0)

#define QMB_CHECK_LAUNCH() QMB_HIP_TRY(hipGetLastError())

#define QMB_TRY(expr)             \
    do {                          \
        int qmb_rc_ = (expr);     \
        if (qmb_rc_ != QMB_OK) {  \
            return qmb_rc_;       \
        }                         \
    } while (0)

// ---------------------------------------------------------------------------------------------
// Workspace carving.  The same routine runs with base == nullptr to size a workspace and with the
// caller's buffer to lay it out, so sizes and layouts cannot drift apart.
// ---------------------------------------------------------------------------------------------
struct Carver {
    char* base = nullptr;
    size_t off = 0;

    template <typename T>
    T* take(size_t count) {
        off = (off + 255) & ~static_cast<size_t>(255);
        T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
        off += count * sizeof(T);
        return p;
    }
    void* take_bytes(size_t bytes) { return take<char>(bytes); }
};

inline int n_words_of(int n_qubytes) { return (n_qubytes + 7) / 8; }

inline uint64_t next_pow2(uint64_t x) {
    uint64_t p = 1;
    while (p < x) {
        p <<= 1;
    }
    return p;
}

inline int check_qubytes(int n_qubytes) {
    if (n_qubytes < 1 || n_qubytes > kMaxQubytes) {
        return set_error(QMB_ERR_UNSUPPORTED, "n_qubytes must be in [1, %d], got %d", kMaxQubytes, n_qubytes);
    }
    return QMB_OK;
}

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Launch geometry of the enumeration kernels.
constexpr int kEnumThreads = 256;

struct LaunchShape {
    int tile_log2;
    int groups;
    bool direct;
    int symmetric;  // 0 off, 1 gather, 2 half-lookup scatter
};

// Term tables up to this size use the direct strategy by default.
constexpr int64_t kDirectMaxTerms = 4096;

LaunchShape resolve_launch(const qmb_launch* launch, int n_words, int64_t n_i, int64_t n_terms);
size_t enum_lds_bytes(int n_words, int n_bits, int tile_log2);
size_t direct_lds_bytes(int n_words);

// ---------------------------------------------------------------------------------------------
// Builders (qmb_build.hip).
// ---------------------------------------------------------------------------------------------
int terms_layout(Carver& cv, int n_qubytes, int64_t n_terms, qmb_terms* out, void** scratch, size_t* scratch_bytes);
// exclude_pairs: terms to leave out (they go to the structured path instead), a mask of:
constexpr int kExcludePairs = 1;     // pair excitations (qmb_pairs)
constexpr int kExcludeSingles = 2;   // single excitations (qmb_singles)
constexpr int kExcludeDiagonal = 4;  // diagonal terms (applied from the diagonal records)
constexpr int kExcludeHops = 8;      // single excitations between two sites of one parity (qmb_apply_hops)
int terms_build(int n_qubytes, int particle_cut, const int16_t* site, const uint8_t* kind, const double* coef,
                int64_t n_terms, void* buffer, size_t bytes, qmb_terms* out, hipStream_t stream, int exclude_pairs = 0);
int targets_layout(Carver& cv, int n_qubytes, int64_t n, qmb_targets* out);
int targets_build(int n_qubytes, const uint8_t* configs, int64_t n, void* buffer, size_t bytes, qmb_targets* out,
                  hipStream_t stream);
// targets_build in two halves: the key words and hash columns, then the hash table, filter and
// duplicate count over them.
int targets_build_keys(int n_qubytes, const uint8_t* configs, int64_t n, void* buffer, size_t bytes, qmb_targets* out,
                       hipStream_t stream);
int targets_build_index(const qmb_targets* t, hipStream_t stream);

// Sorting of config rows by bytes (radix over byte-swapped words).  keys_be: [n][W] output in
// sorted order, perm: original row of each sorted position.
size_t sort_configs_scratch_bytes(int n_words, int64_t n);
int sort_configs(int n_qubytes, const uint8_t* configs, int64_t n, uint64_t* keys_be, uint32_t* perm, void* scratch,
                 size_t scratch_bytes, hipStream_t stream);

// Original-order compiled records (for single_relative), top bit set for never-succeeding terms.
constexpr uint64_t kNeverSucceeds = 1ull << 63;
int compile_records_original(int n_qubytes, int particle_cut, const int16_t* site, const uint8_t* kind, int64_t n_terms,
                             uint64_t* orecs, hipStream_t stream);

// ---------------------------------------------------------------------------------------------
// Kernels launchers (qmb_apply.hip, qmb_find_relative.hip).
// ---------------------------------------------------------------------------------------------
int launch_fill_u64(uint64_t* p, uint64_t value, int64_t n, hipStream_t stream);
int launch_zero_f64(double* p, int64_t n, hipStream_t stream);

int apply_prepared(const qmb_terms* terms, const qmb_targets* targets, const uint8_t* configs_i, int64_t n_i,
                   const double* in, double* out, int mode, const qmb_launch* launch, hipStream_t stream);

int apply_plain(int n_qubytes, int particle_cut, const uint8_t* configs_i, const double* psi_i, int64_t n_i,
                const uint64_t* sorted_keys_be, const uint32_t* perm, int64_t n_j, const int16_t* site, const uint8_t* kind,
                const double* coef, int64_t n_terms, double* psi_j, hipStream_t stream);

int diagonal_prepared(const qmb_terms* terms, const uint8_t* configs, int64_t n, double* out, hipStream_t stream);
int apply_diagonal_terms(const qmb_terms* terms, const qmb_targets* targets, const double* in, double* out, int mode,
                         hipStream_t stream);
int diagonal_values(const qmb_terms* terms, const qmb_targets* targets, double* d, hipStream_t stream);
int apply_diagonal_values(int64_t n, const double* d, const double* in, double* out, int mode, hipStream_t stream);

int single_relative_prepared(int n_qubytes, int particle_cut, const uint8_t* configs, int64_t n, const int16_t* site,
                             const uint8_t* kind, const double* coef, int64_t n_terms, uint64_t seed, const double* keys,
                             const qmb_targets* exclude, uint8_t* out, void* scratch, size_t scratch_bytes,
                             hipStream_t stream);
size_t single_relative_scratch_bytes(int64_t n_terms);

// Pair-excitation path (qmb_pairs.hip).  The dense pair and single tables cover configs of up to
// kPairMaxQubytes bytes (128 sites).
constexpr int kPairMaxQubytes = 16;  // the same bound as the pair tables' layouts
size_t pairs_bytes(int n_qubytes);
int pairs_build(int n_qubytes, int particle_cut, const int16_t* site, const uint8_t* kind, const double* coef, int64_t n_terms,
                void* buffer, size_t bytes, qmb_pairs* out, hipStream_t stream);
int pairs_set_real(qmb_pairs* pairs, int real, hipStream_t stream);
int singles_set_real(qmb_singles* singles, int real, hipStream_t stream);
size_t singles_bytes(int n_qubytes);
int singles_build(int n_qubytes, int particle_cut, const int16_t* site, const uint8_t* kind, const double* coef, int64_t n_terms,
                  void* buffer, size_t bytes, qmb_singles* out, hipStream_t stream);
size_t cores_bytes(int n_qubytes, int64_t n, int max_occupied, int order);
int cores_build(const qmb_targets* targets, int max_occupied, int order, void* buffer, size_t bytes, qmb_cores* out,
                hipStream_t stream);
int apply_pairs(const qmb_pairs* pairs, const qmb_cores* cores, const qmb_targets* targets, const double* in, double* out,
                int mode, hipStream_t stream);
int apply_singles(const qmb_singles* singles, const qmb_cores* cores, const qmb_targets* targets, const double* in, double* out,
                  int mode, hipStream_t stream);
int apply_diagonal(const qmb_pairs* pairs, const qmb_targets* targets, const double* in, double* out, int mode,
                   hipStream_t stream);
size_t strings_bytes(int n_qubytes, int64_t n);
int strings_build(const qmb_targets* targets, int parity, void* buffer, size_t bytes, qmb_strings* out, hipStream_t stream);
int apply_hops(const qmb_singles* singles, const qmb_strings* strings, const qmb_targets* targets, const double* in, double* out,
               int mode, hipStream_t stream);
// Diagonal, single- and pair-excitation parts of A (mode 0) or A^H (mode 1) applied to two vectors
// in one pass each (the forward and backward of a Hermitian training step); accumulates into out
// and out2.
int apply_structured_dual(const qmb_pairs* pairs, const qmb_singles* singles, const qmb_cores* cores2, const qmb_cores* cores1,
                          const qmb_targets* targets, const double* in, double* out, const double* in2, double* out2, int mode,
                          hipStream_t stream);
size_t batch_bytes(int n_items);
int batch_plan(const qmb_batch_item* items, int n_items, void* buffer, size_t bytes, qmb_batch* out, hipStream_t stream);
int apply_batch(const qmb_batch* batch, int mode, int waves, hipStream_t stream);

int sampler_children(int dtype, int64_t n, int site, int n_sites, int spin_up, int spin_down, const uint64_t* prefix, int words,
                     const void* logits, const void* unperturbed, const void* perturbed, const void* uniform, uint64_t seed,
                     void* child_unperturbed, void* child_perturbed, hipStream_t stream);
int sampler_select(int dtype, int64_t n_sel, const int64_t* index, int64_t n_parent, int site, const uint64_t* prefix, int words,
                   const void* child_unperturbed, const void* child_perturbed, uint64_t* out_prefix, void* out_unperturbed,
                   void* out_perturbed, hipStream_t stream);
int sampler_prefix_inputs(int dtype, int64_t n, int site, const uint64_t* prefix, int words, void* out, hipStream_t stream);
int sampler_prefix_configs(int64_t n, int n_sites, const uint64_t* prefix, int words, const int64_t* ordering, int n_qubytes,
                           uint8_t* out, hipStream_t stream);
int network_inputs(int dtype, int64_t n, int n_sites, const uint8_t* configs, int n_qubytes, const int64_t* ordering_reversed, void* out,
                   hipStream_t stream);
int network_amplitude(int dtype, int64_t n, int n_sites, int spin_up, int spin_down, const uint8_t* configs, int n_qubytes,
                      const int64_t* ordering_reversed, const void* logits, double* out, const double* grad, void* grad_logits,
                      hipStream_t stream);

int network_fused_supported(int n_sites, int hidden);
int network_forward_fused(int64_t n, int n_sites, int hidden, int spin_up, int spin_down, const uint8_t* configs, int n_qubytes,
                          const int64_t* ordering_reversed, const float* w1, const float* b1, const float* w2, const float* b2,
                          const float* wp1, const float* bp1, const float* wp2, const float* bp2, double* amplitude, float* phase,
                          hipStream_t stream);
int sampler_step_fused(int64_t n, int site, int n_sites, int hidden, int spin_up, int spin_down, const uint64_t* prefix, int words,
                       const float* w1, const float* b1, const float* w2, const float* b2, const float* unperturbed, const float* perturbed,
                       const float* uniform, uint64_t seed, float* child_unperturbed, float* child_perturbed, hipStream_t stream);

int targets_lookup(const qmb_targets* targets, const uint8_t* configs, int64_t n, int64_t* rows, hipStream_t stream);

size_t lanczos_bytes();
int lanczos_step(int64_t n, const double* v, const double* v_prev, double beta, double* w, double* v_next, double* scalars,
                 void* workspace, size_t bytes, hipStream_t stream);

size_t find_relative_scratch_bytes(int n_qubytes, int64_t n_i, int64_t count_selected, int64_t n_terms);
int find_relative_prepared(const qmb_terms* terms, const qmb_targets* exclude, const uint8_t* configs_i, const double* psi_i,
                           int64_t n_i, int64_t count_selected, uint8_t* out, int64_t* out_rows, void* scratch,
                           size_t scratch_bytes, hipStream_t stream);

}  // namespace qmb
