/*
 * qmb_hip.h -- C ABI of libqmb_hip.so, the MI355X (gfx950) implementation of qmb's sparse
 * second-quantized Hamiltonian operators on bit-packed configurations.
 *
 * Conventions shared by every entry point:
 *   - All array arguments are DEVICE pointers (hipMalloc / torch HIP tensors), row-major, contiguous.
 *   - A config is n_qubytes bytes; site s is bit (s % 8) of byte s / 8 (qmb/bitspack.py:50).
 *   - A term table is site int16[n_terms][4], kind uint8[n_terms][4] (0 = a, 1 = a^dagger,
 *     2 = identity pad), coef float64[n_terms][2] (re, im) -- the layout of qmb/_hamiltonian.cpp:29-64.
 *   - Amplitudes are complex128 stored as float64 pairs (re, im).
 *   - particle_cut 1 = fermions (Jordan-Wigner sign), 2 = hard-core bosons (no sign).
 *   - n_qubytes is 1..32.
 *   - `stream` is a hipStream_t passed as void* (0 = the null stream).  Every call only enqueues
 *     work on that stream and returns; no entry point synchronises the device or allocates device
 *     memory.  Scratch memory comes from the caller through `workspace`, whose size the matching
 *     *_workspace_size function returns.
 *   - Return value: QMB_OK or an error code; qmb_last_error() gives the message of the last failure
 *     on the calling host thread.
 *   - Preconditions the reference also assumes: target / exclude config sets are duplicate-free
 *     (with duplicates, the lowest row index receives the contribution).
 */
#ifndef QMB_HIP_H
#define QMB_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QMB_OK 0
#define QMB_ERR_ARGUMENT 1
#define QMB_ERR_HIP 2
#define QMB_ERR_WORKSPACE 3
#define QMB_ERR_UNSUPPORTED 4

/* Message of the last failing call on this host thread ("" if none). */
const char* qmb_last_error(void);

/* Library version as major * 10000 + minor * 100 + patch. */
int qmb_version(void);

/* Host-only validation of a term table held in HOST memory: counts terms whose kind is outside
 * {0,1,2} or whose site lies outside [0, 8 * n_qubytes).  Replaces nothing in the reference
 * (which leaves both undefined, qmb/_hamiltonian.cpp:29-64); the Python layer calls it once per
 * table and n_qubytes. */
int qmb_validate_terms_host(int n_qubytes, const int16_t* site, const uint8_t* kind, int64_t n_terms, int64_t* n_bad);

/* Host-only compilation of one term into the library's 64-bit record (for unit tests of the
 * term semantics without a GPU).  status: 0 ok, 1 never succeeds, 2 bad site, 3 bad kind. */
int qmb_compile_term_host(const int16_t* site4, const uint8_t* kind4, int particle_cut, int n_qubytes, uint64_t* rec,
                          int32_t* status, int64_t* bucket, int32_t* diagonal);

/* Host-only application of a compiled record to one config (n_qubytes bytes, host memory).
 * Returns 1 on success and writes the resulting config and parity bit. */
int qmb_apply_record_host(uint64_t rec, int n_qubytes, const uint8_t* config, uint8_t* result, int32_t* parity);

/* Host-only Hermiticity test of a term table held in HOST memory: *hermitian = 1 when the operator
 * sum_t coef_t O_t equals its adjoint, comparing the compiled form of every term with that of its
 * adjoint (terms with the same action are summed first) to a relative tolerance `tol`. */
int qmb_terms_hermitian_host(int n_qubytes, int particle_cut, const int16_t* site, const uint8_t* kind, const double* coef,
                             int64_t n_terms, double tol, int32_t* hermitian);

/* Host-only Philox4x32-10 block (for known-answer tests) and the (0,1] uniform built on it. */
void qmb_philox4x32_10_host(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double qmb_philox_uniform_host(uint64_t seed, uint64_t index);

/* ============================== drop-in operators ============================================ */

/* apply_within: psi_j[k] = sum over (i, t) with t(configs_i[i]) == configs_j[k] of
 * sign * coef[t] * psi_i[i]; psi_j is written in configs_j's own row order.
 * Replaces apply_within_interface, qmb/_hamiltonian_cuda.cu:171-274 (CPU: _hamiltonian_cpu.cpp:164-257),
 * bound by qmb/hamiltonian.py:104-131. */
/* Path: when configs_i and configs_j are the same buffer (same pointer, same row count) and the
 * workspace is at least qmb_apply_within_structured_workspace_size for the set, the call reads four
 * counters back from the device (one synchronisation of `stream`, as the reference's op
 * synchronises after every call) and, when the table has pair excitations, the rows are distinct and
 * the workspace holds the core indexes for the set's most occupied row, applies the structured path
 * (pair and single excitations by core groups, the folded diagonal, the remaining terms enumerated).
 * Failing that, with a workspace of at least qmb_apply_within_hops_workspace_size, a table with
 * hops between sites of one parity and distinct rows take the hop path (rows grouped by the other
 * parity's string, two more synchronisations: the term classes, then the pair-test count, which
 * must stay within 4 per row and hop term).  Otherwise -- two config sets, or a workspace of
 * qmb_apply_within_workspace_size -- it enumerates every connection with hash lookups and does
 * not synchronise.  All give the same results (to FP64 summation order). */
int qmb_apply_within_workspace_size(int n_qubytes, int64_t n_i, int64_t n_j, int64_t n_terms, size_t* bytes);
/* Workspace for the structured path on one set of n configs whose rows hold at most max_occupied
 * occupied sites (e.g. the particle number); n_qubytes <= 16.  No reference counterpart. */
int qmb_apply_within_structured_workspace_size(int n_qubytes, int64_t n, int64_t n_terms, int32_t max_occupied, size_t* bytes);
/* Workspace for the hop path on one set of n configs; n_qubytes <= 16.  No reference counterpart. */
int qmb_apply_within_hops_workspace_size(int n_qubytes, int64_t n, int64_t n_terms, size_t* bytes);
int qmb_apply_within(int n_qubytes, int particle_cut, const uint8_t* configs_i, const double* psi_i, int64_t n_i,
                     const uint8_t* configs_j, int64_t n_j, const int16_t* site, const uint8_t* kind, const double* coef,
                     int64_t n_terms, double* psi_j, void* workspace, size_t workspace_bytes, void* stream);

/* find_relative: the configs reached from configs_i that are not in configs_exclude, ranked per
 * contribution by |sign * coef[t] * psi_i[i]|^2; the count_selected largest contributions are kept,
 * zero weights dropped, and the distinct configs among them are written sorted by bytes into
 * out (capacity count_selected rows).  *out_rows (device int64) receives the row count.
 * Entries tied with the count_selected-th weight are taken in ascending config order.
 * Replaces find_relative_interface, qmb/_hamiltonian_cuda.cu:551-675 (CPU :442-546),
 * bound by qmb/hamiltonian.py:133-166. */
int qmb_find_relative_workspace_size(int n_qubytes, int64_t n_i, int64_t n_exclude, int64_t n_terms, int64_t count_selected,
                                     size_t* bytes);
int qmb_find_relative(int n_qubytes, int particle_cut, const uint8_t* configs_i, const double* psi_i, int64_t n_i,
                      int64_t count_selected, const int16_t* site, const uint8_t* kind, const double* coef, int64_t n_terms,
                      const uint8_t* configs_exclude, int64_t n_exclude, uint8_t* out, int64_t* out_rows, void* workspace,
                      size_t workspace_bytes, void* stream);

/* diagonal_term: out[k] = sum over terms t with t(configs[k]) == configs[k] of sign * coef[t].
 * Replaces diagonal_term_interface, qmb/_hamiltonian_cuda.cu:738-803, bound by
 * qmb/hamiltonian.py:168-186. */
int qmb_diagonal_term_workspace_size(int n_qubytes, int64_t n, int64_t n_terms, size_t* bytes);
int qmb_diagonal_term(int n_qubytes, int particle_cut, const uint8_t* configs, int64_t n, const int16_t* site,
                      const uint8_t* kind, const double* coef, int64_t n_terms, double* out, void* workspace,
                      size_t workspace_bytes, void* stream);

/* single_relative: for each config, among terms whose result is not one of the input configs, the
 * result of the term with the largest key u_t^(1/|coef_t|) (Efraimidis-Spirakis), u_t uniform in
 * (0,1] from Philox4x32-10 keyed by `seed` at counter t; an all-zero row when there is none.
 * Ties go to the lower term index.  Replaces single_relative_interface,
 * qmb/_hamiltonian_cuda.cu:907-996 (whose keys come from curand), bound by
 * qmb/hamiltonian.py:188-206.  If `keys` is non-NULL it supplies the n_terms keys (device
 * float64) instead of the Philox draw. */
int qmb_single_relative_workspace_size(int n_qubytes, int64_t n, int64_t n_terms, size_t* bytes);
int qmb_single_relative(int n_qubytes, int particle_cut, const uint8_t* configs, int64_t n, const int16_t* site,
                        const uint8_t* kind, const double* coef, int64_t n_terms, uint64_t seed, const double* keys,
                        uint8_t* out, void* workspace, size_t workspace_bytes, void* stream);

/* ============================== extensions (no reference counterpart) ========================= */

/* Adjoint of apply_within with respect to psi_i: grad_i[i] = sum over (t, k) with
 * t(configs_i[i]) == configs_j[k] of conj(sign * coef[t]) * grad_j[k] -- a gather, no atomics on
 * the output rows of other configs.  New functionality: the reference never differentiates the
 * path (qmb/vmc.py:104-116 calls apply_within under no_grad). */
int qmb_apply_within_adjoint(int n_qubytes, int particle_cut, const uint8_t* configs_i, int64_t n_i,
                             const uint8_t* configs_j, const double* grad_j, int64_t n_j, const int16_t* site,
                             const uint8_t* kind, const double* coef, int64_t n_terms, double* grad_i, void* workspace,
                             size_t workspace_bytes, void* stream);

/* apply_within without a workspace, for callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  Same arguments and results as qmb_apply_within, minus the workspace: the term table
 * and the target index are built in a fixed device array of qmb_static_scratch_bytes() (64 MiB)
 * that the library holds on every device it runs on until the process ends.  Calls whose
 * term table and target index need more (about 50 bytes per term plus 40 per target config)
 * fail with QMB_ERR_WORKSPACE.  The call allocates nothing and enqueues only.  All calls on one
 * device share that array: calls that overlap in time on one device (different streams, or
 * threads) race and must be ordered by the caller. */
size_t qmb_static_scratch_bytes(void);
int qmb_apply_within_noalloc(int n_qubytes, int particle_cut, const uint8_t* configs_i, const double* psi_i, int64_t n_i,
                             const uint8_t* configs_j, int64_t n_j, const int16_t* site, const uint8_t* kind, const double* coef,
                             int64_t n_terms, double* psi_j, void* stream);

/* The reference's own scheme on the GPU, kept as the comparator: one thread per (term, config),
 * a lexicographic sort of configs_j, a binary search per connection and FP64 atomics
 * (qmb/_hamiltonian_cuda.cu:86-169 restated).  Same results as qmb_apply_within. */
int qmb_apply_within_plain_workspace_size(int n_qubytes, int64_t n_j, size_t* bytes);
int qmb_apply_within_plain(int n_qubytes, int particle_cut, const uint8_t* configs_i, const double* psi_i, int64_t n_i,
                           const uint8_t* configs_j, int64_t n_j, const int16_t* site, const uint8_t* kind,
                           const double* coef, int64_t n_terms, double* psi_j, void* workspace, size_t workspace_bytes,
                           void* stream);

/* ---------------- prepared objects: build once, apply many times (Krylov loops, autograd) ------ */

/* A compiled, bucketed term table living in a caller-owned device buffer. */
typedef struct qmb_terms {
    int32_t n_qubytes, particle_cut, n_words, n_bits;
    int64_t n_terms, n_buckets;
    void* buffer;      /* device buffer the fields below point into */
    size_t bytes;
    uint64_t* recs;    /* [n_terms] compiled records, bucket order (valid prefix: counts[0]) */
    double* coefs;     /* [n_terms][2] matching coefficients */
    uint32_t* tidx;    /* [n_terms] original term index */
    uint32_t* bucket_start; /* [n_buckets + 1] */
    uint64_t* drecs;   /* [n_terms] diagonal terms in original order (valid prefix: counts[1]) */
    double* dcoefs;    /* [n_terms][2] */
    int64_t* counts;   /* device: {valid terms, diagonal terms, invalid terms, 0} */
} qmb_terms;

int qmb_terms_workspace_size(int n_qubytes, int64_t n_terms, size_t* bytes);
int qmb_terms_build(int n_qubytes, int particle_cut, const int16_t* site, const uint8_t* kind, const double* coef,
                    int64_t n_terms, void* buffer, size_t bytes, qmb_terms* out, void* stream);

/* A hash index over a config set: padded key words, an open-addressing table and a blocked Bloom
 * filter (16 bits per key) that answers most misses from L2 before the table is touched. */
typedef struct qmb_targets {
    int32_t n_qubytes, n_words;
    int64_t n;
    uint64_t capacity;     /* power of two, >= 4 n */
    uint64_t filter_words; /* power of two, ~ n / 4 */
    void* buffer;
    size_t bytes;
    uint64_t* keys;    /* [n][n_words] */
    uint64_t* table;   /* [capacity] (index << 32 | fingerprint) or ~0 */
    uint64_t* filter;  /* [filter_words] */
    uint64_t* cols;    /* [256] hash column per site: H(config) = XOR of cols over occupied sites */
    int64_t* duplicates; /* device: number of rows equal to an earlier row */
} qmb_targets;

int qmb_targets_workspace_size(int n_qubytes, int64_t n, size_t* bytes);
/* rows[i] = the row of configs[i] (n_qubytes bytes) in the indexed set (the lowest one when the set
 * repeats it), or -1.  No counterpart in the reference: it is how an application across two
 * config sets is mapped onto one set (see qmb_hip.prepared.apply_across). */
int qmb_targets_lookup(const qmb_targets* targets, const uint8_t* configs, int64_t n, int64_t* rows, void* stream);
int qmb_targets_build(int n_qubytes, const uint8_t* configs, int64_t n, void* buffer, size_t bytes, qmb_targets* out,
                      void* stream);
/* qmb_targets_build in two halves, so a caller whose kernels read only the key words and hash
 * columns (the core, strings and hop paths) need not build the hash index: _keys lays out the
 * whole descriptor and fills keys and cols; _index then fills table, filter and duplicates in the
 * same buffer.  No reference counterpart. */
int qmb_targets_build_keys(int n_qubytes, const uint8_t* configs, int64_t n, void* buffer, size_t bytes, qmb_targets* out,
                           void* stream);
int qmb_targets_build_index(const qmb_targets* targets, void* stream);

/* Launch shape for the enumeration kernels (0 = library default). */
typedef struct qmb_launch {
    int32_t tile;     /* bucketed strategy: configs per workgroup tile, 64..4096, power of two */
    int32_t groups;   /* bucketed strategy: bucket groups per tile (bucket b goes to group b % groups) */
    int32_t strategy; /* 0 auto (direct for <= 4096 terms), 1 bucketed tiles, 2 direct (every config
                         walks every term from LDS) */
    int32_t symmetric; /* nonzero: the caller asserts the operator is Hermitian (see
                          qmb_terms_hermitian_host) and configs_i is the indexed set itself with no
                          repeated rows, so modes 0 and 1 compute the same map (A^H = A).
                          1: both run as the gather of mode 1 (per-config sums, no atomics per
                          contribution); 2: each connected pair is looked up once and both A_ji and
                          conj(A_ji) are scattered with atomics. */
} qmb_launch;

/* mode 0: forward scatter  out[k] += sum sign*coef*in[i]      (in = psi_i [n_i], out = psi_j [n_j])
 * mode 1: adjoint gather   out[i] += sum conj(sign*coef)*in[k] (in = grad_j [n_j], out = grad_i [n_i])
 * mode 2: count            ((uint64_t*)out)[0] += connections (successful term applications),
 *                          ((uint64_t*)out)[1] += contributions (connections whose target is indexed);
 *                          `in` is unused.  These are M_gen and M_hit of the traffic model.
 * `out` is accumulated into (zero it first).  configs_i are n_qubytes-byte rows. */
int qmb_apply_prepared(const qmb_terms* terms, const qmb_targets* targets, const uint8_t* configs_i, int64_t n_i,
                       const double* in, double* out, int mode, const qmb_launch* launch, void* stream);

/* ---------------- pair-excitation path (no reference counterpart) ------------------------------
 * A pair excitation is a term on four distinct sites that annihilates two and creates two, with
 * no condition on other sites (the two-body integrals of a molecular Hamiltonian).  Instead of
 * generating every (config, term) target and looking it up, the pair path groups the configs of
 * one set by "core" (the config with two occupied sites removed): two configs of the set are
 * connected by a pair excitation exactly when they share a core with disjoint removed pairs, so
 * only connections whose target is in the set are ever visited.  The remaining terms stay on the
 * enumeration path (qmb_terms_build_residual).  Applies within one config set (configs_i indexed
 * by the qmb_targets the cores were built from, rows distinct). */

/* Merged coefficients of the pair-excitation terms, dense over (annihilated pair, created pair):
 * coef[pair_slot(r0,r1) * n_pairs + pair_slot(c0,c1)] = sum over terms of (-1)^delta_t coef_t.
 * pair_slot numbers the pairs (a<b) of the n_bits = 8*n_qubytes sites in three blocks by site
 * parity: even-even pairs (index (b/2)((b/2)-1)/2 + a/2), then odd-odd, then even-odd
 * (2 C(n_bits/2, 2) + (even site / 2) (n_bits / 2) + odd site / 2).  n_qubytes <= 16.
 * Also the diagonal terms, folded exactly (a diagonal term conditions at most two sites and has a
 * constant sign) into diag = {constant, per site p, per site pair (p, q)}: the diagonal element of
 * a config c is diag[0] + sum_{p in c} diag[1 + p] + sum_{p<q in c} diag[1 + n_bits + pair_index(p, q)]. */
typedef struct qmb_pairs {
    int32_t n_qubytes, particle_cut, n_bits;
    int32_t real;      /* 1: the kernels read coef_re (set by qmb_pairs_set_real) */
    int64_t n_pairs;
    void* buffer;
    size_t bytes;
    double* coef;      /* [n_pairs * n_pairs][2] */
    double* coef_re;   /* [n_pairs * n_pairs] real parts, filled by qmb_pairs_set_real */
    double* diag;      /* [1 + n_bits + n_pairs][2] */
    int64_t* counts;   /* device: {pair-excitation terms, diagonal terms, entries with a nonzero imaginary
                          part (after qmb_pairs_set_real), 0} */
} qmb_pairs;

int qmb_pairs_workspace_size(int n_qubytes, size_t* bytes);
int qmb_pairs_build(int n_qubytes, int particle_cut, const int16_t* site, const uint8_t* kind, const double* coef,
                    int64_t n_terms, void* buffer, size_t bytes, qmb_pairs* out, void* stream);

/* Merged coefficients of the single excitations (annihilate r, create p, conditioned on at most
 * one other site k): the element on a source config c holding r and not p is
 *   (-1)^(popcount(c & (L(r) ^ L(p)))) * (a[r * n_bits + p] + sum_{k in c, k != r} kd[(r * n_bits + p) * n_bits + k])
 * with L(s) = the sites below s.  n_qubytes <= 16. */
typedef struct qmb_singles {
    int32_t n_qubytes, particle_cut, n_bits;
    int32_t real;      /* 1: the kernels read a_re / kd_re (set by qmb_singles_set_real) */
    void* buffer;
    size_t bytes;
    double* a;         /* [n_bits * n_bits][2] */
    double* kd;        /* [n_bits * n_bits * n_bits][2] */
    int64_t* counts;   /* device: {single-excitation terms, those with a conditioned site (c+_p n_k c_q),
                          entries with a nonzero imaginary part (after qmb_singles_set_real), 0} */
    double* a_re;      /* [n_bits * n_bits] real parts, filled by qmb_singles_set_real */
    double* kd_re;     /* [n_bits * n_bits * n_bits] */
} qmb_singles;

int qmb_singles_workspace_size(int n_qubytes, size_t* bytes);
int qmb_singles_build(int n_qubytes, int particle_cut, const int16_t* site, const uint8_t* kind, const double* coef,
                      int64_t n_terms, void* buffer, size_t bytes, qmb_singles* out, void* stream);

/* For tables whose coefficients are all real: copies the real parts to coef_re and makes the
 * kernels read those (half the bytes per coefficient).  real = 0 switches back to the complex
 * table.  Counts the entries with a nonzero imaginary part into counts[2]; the caller asserts
 * there are none. */
int qmb_pairs_set_real(qmb_pairs* pairs, int32_t real, void* stream);
/* The same for the single-excitation table (a and kd).  No reference counterpart. */
int qmb_singles_set_real(qmb_singles* singles, int32_t real, void* stream);

/* qmb_terms_build without some terms (same layout and buffer size).  exclude is a mask:
 * 1 pair excitations, 2 single excitations, 4 diagonal terms (from the buckets; the diagonal
 * records are kept), 8 single excitations between two sites of one parity (the hop path). */
int qmb_terms_build_residual(int n_qubytes, int particle_cut, const int16_t* site, const uint8_t* kind, const double* coef,
                             int64_t n_terms, int32_t exclude, void* buffer, size_t bytes, qmb_terms* out, void* stream);

/* Core index of the set indexed by a qmb_targets: one entry per (config, `order` occupied sites)
 * -- order 2: every occupied pair (cores of pair excitations), order 1: every occupied site
 * (cores of single excitations) -- radix-sorted by a 32-bit key of the core (its class, order 2:
 * even sites of the core mod 3, order 1: 0, and a hash of it), stable within a group of equal keys;
 * and the start of every group.  Different cores with equal keys share a group: the build compares
 * the cores of neighbouring entries exactly and flags such a mixed group in bit 31 of its start,
 * and the apply calls compare cores per pair only there (elsewhere every row of a group has the
 * first entry's core).  capacity must be below 2^31.  max_occupied must bound the number of
 * occupied sites of every config (a row with more sets counts[2], and the apply calls then write
 * NaN into out[0] instead of a wrong result). */
typedef struct qmb_cores {
    int32_t n_qubytes, n_words, max_occupied, per_config; /* per_config = C(max_occupied, order) */
    int32_t order, key_bits;  /* key_bits: width of the sorted keys, 32 (set by qmb_cores_build) */
    int64_t n;
    int64_t capacity;  /* n * per_config entries */
    void* buffer;
    size_t bytes;
    uint32_t* keys;    /* [capacity] sorted keys: core class (even sites of the core mod 3) * 0x55555555 + the
                          64-bit core hash mod 0x55555555; a group is a run of equal keys; unused entries last
                          (0xFFFFFFFF) */
    uint64_t* vals;    /* [capacity] row << 16 | site1 << 8 | site0 of the removed sites (order 1: site0) */
    uint32_t* starts;  /* [capacity + 1] first entry of each group (bit 31: the group is mixed), then the end */
    int64_t* counts;   /* device: {entries, groups, overflow, 0} */
} qmb_cores;

int qmb_cores_workspace_size(int n_qubytes, int64_t n, int32_t max_occupied, int32_t order, size_t* bytes);
int qmb_cores_build(const qmb_targets* targets, int32_t max_occupied, int32_t order, void* buffer, size_t bytes, qmb_cores* out,
                    void* stream);

/* Adds the pair-excitation part of mode 0 (forward, out = A in) or mode 1 (adjoint,
 * out = A^H in) on the set of `cores`/`targets` into out (float64 [n][2]).  A gather over the
 * core groups: one FP64 atomic per (config, group) sum, none per contribution. */
int qmb_apply_pairs(const qmb_pairs* pairs, const qmb_cores* cores, const qmb_targets* targets, const double* in, double* out,
                    int mode, void* stream);

/* The single-excitation part (order-1 cores), as qmb_apply_pairs. */
int qmb_apply_singles(const qmb_singles* singles, const qmb_cores* cores, const qmb_targets* targets, const double* in,
                      double* out, int mode, void* stream);

/* The diagonal part on the set indexed by `targets`: out[i] += d_i in[i] (mode 0) or
 * conj(d_i) in[i] (mode 1), d_i from the folded diagonal table of `pairs`. */
int qmb_apply_diagonal(const qmb_pairs* pairs, const qmb_targets* targets, const double* in, double* out, int mode, void* stream);

/* Hop path, for tables whose off-diagonal terms are single excitations between two sites of one
 * parity (spin-conserving hops when sites interleave spin, even = up, odd = down: the Hubbard model
 * of qmb/hubbard.py).  No reference counterpart: the reference enumerates every
 * term on every config and binary-searches the result (qmb/_hamiltonian_cuda.cu:86-169); here a
 * hop connects two rows of the set only if they hold the same string on the other parity's sites,
 * so the rows are grouped by that string and the rows of a group are tested pairwise.
 *
 * The strings index of parity s (0: hops between even sites, 1: between odd sites): one entry per
 * row, radix-sorted by a 32-bit key of the row's string on the sites of parity 1 - s (a 64-bit
 * hash of it mod 2^32 - 1; rows of different strings with equal keys share a group, flagged as
 * merged, and are told apart per pair). */
typedef struct qmb_strings {
    int32_t n_qubytes, n_words, parity, pad;
    int64_t n;
    void* buffer;
    size_t bytes;
    uint32_t* keys;    /* [n] sorted keys */
    uint32_t* rows;    /* [n] row of each entry */
    uint32_t* group;   /* [n] group of each entry */
    uint32_t* starts;  /* [n + 1] first entry of each group, then the end */
    /* The entries of the groups of two or more rows (the only ones qmb_apply_hops walks), in index
     * order: */
    uint32_t* mrows;   /* [n] their rows */
    uint32_t* mrange;  /* [n][2] their group as [first, end) of this order; bit 31 of the end marks a
                          group whose rows hold more than one other-parity string */
    uint64_t* mcfg;    /* [n][n_words] their config words */
    uint64_t* mstr;    /* [n] their hop-parity string: bit j = site 2 j + parity */
    int64_t* counts;   /* device: {entries, groups, sum of size^2 over the groups of two or more rows (the
                          pairs qmb_apply_hops tests), largest group, entries of those groups, 0, 0, 0} */
} qmb_strings;

int qmb_strings_workspace_size(int n_qubytes, int64_t n, size_t* bytes);
int qmb_strings_build(const qmb_targets* targets, int32_t parity, void* buffer, size_t bytes, qmb_strings* out, void* stream);

/* Adds the hops of parity strings->parity of mode 0 (out = A in) or mode 1 (out = A^H in) on the
 * set: for each pair of rows of a group that differ in two sites of that parity, one occupied in
 * each, the single-excitation element of `singles` between them (with its conditioned-site terms)
 * and the fermionic sign (particle_cut 1).  Each row is written by one lane, without atomics. */
int qmb_apply_hops(const qmb_singles* singles, const qmb_strings* strings, const qmb_targets* targets, const double* in,
                   double* out, int mode, void* stream);

/* The diagonal terms of `terms` (its diagonal records) on the set indexed by `targets`: out[i] +=
 * d_i in[i] (mode 0) or conj(d_i) in[i] (mode 1), d_i summed over the diagonal terms as
 * qmb_diagonal_term does (qmb/_hamiltonian_cuda.cu:677-708).  For tables with few diagonal terms,
 * where qmb_apply_diagonal's folded table costs a read per occupied pair. */
int qmb_apply_diagonal_terms(const qmb_terms* terms, const qmb_targets* targets, const double* in, double* out, int mode,
                             void* stream);

/* The same diagonal in two halves, for a set applied to more than once (the forward and backward of
 * a step, a Krylov loop): d[i] (float64 [n][2]) = the diagonal value of row i, then out[i] +=
 * d[i] in[i] (mode 0) or conj(d[i]) in[i] (mode 1) per application.  No reference counterpart. */
int qmb_diagonal_values(const qmb_terms* terms, const qmb_targets* targets, double* d, void* stream);
int qmb_apply_diagonal_values(int64_t n, const double* d, const double* in, double* out, int mode, void* stream);

/* The diagonal, single- and pair-excitation parts of A (mode 0) or A^H (mode 1) on two vectors at
 * once, for one set: out += A in and out2 += A in2, each kernel walking its groups once for both
 * (for a Hermitian table, the forward and the backward of a training step).  singles/cores1 may be
 * NULL for a table without single excitations.  The per-instance form of a two-vector batch.  No
 * reference counterpart. */
int qmb_apply_structured_dual(const qmb_pairs* pairs, const qmb_singles* singles, const qmb_cores* cores2, const qmb_cores* cores1,
                              const qmb_targets* targets, const double* in, double* out, const double* in2, double* out2, int mode,
                              void* stream);

/* ---------------- batched persistent launch (no reference counterpart) -------------------------
 * Many instances -- a term table and a config set each, of any sizes -- applied by two persistent
 * kernel launches for all of them: the diagonal, single- and pair-excitation work of every instance
 * is cut into units of 64 core groups or 64 rows; the first launch's waves take the pair units of
 * all instances round-robin, the second's the single and diagonal units.  For many small instances
 * this replaces a few kernels per instance by two launches that fill the GPU; both are capturable
 * in a hipGraph (no allocation, no sync).
 * An instance whose table has terms outside those classes (its residual table) needs that part
 * applied separately. */
typedef struct qmb_batch_item {
    const qmb_terms* terms;      /* the full table (kept for the caller's residual part; unused here) */
    const qmb_pairs* pairs;      /* pair-excitation and diagonal tables (required) */
    const qmb_singles* singles;  /* single-excitation table or NULL */
    const qmb_targets* targets;  /* index of the instance's config set (n_qubytes <= 16) */
    const qmb_cores* cores2;     /* order-2 core index of that set (with pairs) */
    const qmb_cores* cores1;     /* order-1 core index of that set (with singles) */
    const double* in;            /* [n][2] */
    double* out;                 /* [n][2], accumulated into */
    const double* in2;           /* two-vector batches: a second input [n][2] (NULL otherwise) */
    double* out2;                /* and its output, accumulated into */
} qmb_batch_item;

typedef struct qmb_batch {
    int32_t n_items;
    int32_t dual;      /* 1: every item has in2/out2 */
    int64_t total_units;
    void* buffer;
    size_t bytes;
    void* items;       /* device table of the items */
    int64_t* offsets;  /* device [3 * n_items + 1]: first unit of (type, item) */
} qmb_batch;

int qmb_batch_workspace_size(int n_items, size_t* bytes);
/* Copies the item table to the device; synchronises `stream` once (set-up, not per launch). */
int qmb_batch_plan(const qmb_batch_item* items, int n_items, void* buffer, size_t bytes, qmb_batch* out, void* stream);
/* Two persistent launches of `waves` wavefronts each (0: 8192) over all units (the pair units, then
 * the single and diagonal units); mode 0 forward, 1 adjoint
 * (a Hermitian instance gives the same map in both).  A two-vector batch (dual) computes
 * out = A in and out2 = A in2 (mode 0; A^H with mode 1) in the same pass, sharing every group walk
 * and coefficient read: for Hermitian instances that is the forward and the backward of a
 * training step in one pass. */
int qmb_apply_batch(const qmb_batch* batch, int mode, int waves, void* stream);

/* ---------------------------------------------------------------------------------------------
 * Lanczos vector step.  Replaces the tensor arithmetic between two applications in the Krylov
 * loops of qmb/haar.py:123-142 (_DynamicLanczos._run) and qmb/rldiag.py:23-44 (lanczos_energy):
 *     alpha = Re(w^H v);  w -= alpha v + beta v_prev;  |w|;  v_next = w / |w|
 * on complex128 vectors stored as float64 [n][2].  On entry w = H v; v_prev may be null (the first
 * step: no beta term).  Writes scalars[0] = alpha and scalars[1] = |w| (device doubles), updates w
 * in place and writes v_next (zeros when |w| = 0).  Deterministic: fixed-order reductions.  The
 * workspace is a fixed 16 KiB of partial sums; calls that share one workspace must be ordered.
 * ------------------------------------------------------------------------------------------- */
int qmb_lanczos_workspace_size(size_t* bytes);
int qmb_lanczos_step(int64_t n, const double* v, const double* v_prev, double beta, double* w, double* v_next, double* scalars,
                     void* workspace, size_t bytes, void* stream);

/* ---------------------------------------------------------------------------------------------
 * Autoregressive electron up/down network (qmb/mlp.py:61-314, WaveFunctionElectronUpDown): the
 * element-wise and bit-level parts of its forward and of its unique sampler; the GEMMs of the
 * network run in the caller's BLAS.  dtype 0 = float32, 1 = float64 (the network's dtype).
 * Sites are taken in sampling order: site k is physical site ordering[k] (config bits
 * 2 ordering[k], 2 ordering[k] + 1 = spin up, spin down).  A sampled prefix is a bit string of two
 * bits per site in sampling order, in `words` uint64 words per row.  Logits are site-major and
 * row-minor: [4][n] per sampling step, [S][4][n] for the forward; child / entry c = 2 up + down.
 * ------------------------------------------------------------------------------------------- */
/* One sampling step at `site` (qmb/mlp.py:247-270): masks the children a prefix can no longer
 * take, normalises the logits (sum exp(2 x) = 1), and writes for child c of row r, at c n + r,
 * its log-probability (unperturbed[r] + 2 x[c]) and its Gumbel-perturbed value conditioned on the
 * parent's perturbed[r].  The uniforms come from `uniform` ([4][n], values in (0, 1)) when given,
 * else from Philox4x32-10 under `seed`. */
int qmb_sampler_children(int dtype, int64_t n, int site, int n_sites, int spin_up, int spin_down, const uint64_t* prefix, int words,
                         const void* logits, const void* unperturbed, const void* perturbed, const void* uniform, uint64_t seed,
                         void* child_unperturbed, void* child_perturbed, void* stream);
/* Keeps the children index[0..n_sel) (index = c n_parent + r): their prefixes (parent's plus the
 * child's two bits at `site`) and their two log-probabilities (qmb/mlp.py:272-287). */
int qmb_sampler_select(int dtype, int64_t n_sel, const int64_t* index, int64_t n_parent, int site, const uint64_t* prefix, int words,
                       const void* child_unperturbed, const void* child_perturbed, uint64_t* out_prefix, void* out_unperturbed,
                       void* out_perturbed, void* stream);
/* The network input of a sampling step: the first 2 site prefix bits as 0/1 values, [2 site][n]. */
int qmb_sampler_prefix_inputs(int dtype, int64_t n, int site, const uint64_t* prefix, int words, void* out, void* stream);
/* Full prefixes -> packed configs of n_qubytes bytes (qmb/mlp.py:289-295, ordering a permutation). */
int qmb_sampler_prefix_configs(int64_t n, int n_sites, const uint64_t* prefix, int words, const int64_t* ordering, int n_qubytes,
                               uint8_t* out, void* stream);
/* The network input of the forward (qmb/mlp.py:177-181): config bits in sampling order, [2 S][n]. */
int qmb_network_inputs(int dtype, int64_t n, int n_sites, const uint8_t* configs, int n_qubytes, const int64_t* ordering_reversed,
                       void* out, void* stream);
/* The forward's log-amplitude (qmb/mlp.py:183-199): out[r] = sum over sites (in order, double) of
 * the selected normalised logit.  With grad/grad_logits (both or neither) it writes instead
 * grad_logits[k][c][r] = (delta(c, selected) - exp(2 x[c])) grad[r], the gradient of that sum. */
int qmb_network_amplitude(int dtype, int64_t n, int n_sites, int spin_up, int spin_down, const uint8_t* configs, int n_qubytes,
                          const int64_t* ordering_reversed, const void* logits, double* out, const double* grad, void* grad_logits,
                          void* stream);

/* Fused one-hidden-layer float32 networks (hidden width 64, 128, 256 or 512; 2 n_sites hidden
 * columns of float32 must fit LDS): returns 1 when the two calls below support (n_sites, hidden).
 * Weights, float32, row-major: w1 = the site networks' first-layer weights transposed and
 * concatenated (site i: [2 i][hidden] at row i (i - 1)), b1 [S][hidden], w2 [S][4][hidden],
 * b2 [S][4]; the phase network wp1 [2 S][hidden] (transposed), bp1 [hidden], wp2 [hidden], bp2 [1]. */
int qmb_network_fused_supported(int n_sites, int hidden);
/* The whole forward (qmb/mlp.py:169-205) without materialising activations: amplitude[r] = the
 * double log-amplitude, phase[r] = the phase network's output. */
int qmb_network_forward_fused(int64_t n, int n_sites, int hidden, int spin_up, int spin_down, const uint8_t* configs, int n_qubytes,
                              const int64_t* ordering_reversed, const float* w1, const float* b1, const float* w2, const float* b2,
                              const float* wp1, const float* bp1, const float* wp2, const float* bp2, double* amplitude, float* phase,
                              void* stream);
/* qmb_sampler_children with the site network evaluated in the same kernel (prefix words <= 2). */
int qmb_sampler_step_fused(int64_t n, int site, int n_sites, int hidden, int spin_up, int spin_down, const uint64_t* prefix, int words,
                           const float* w1, const float* b1, const float* w2, const float* b2, const float* unperturbed,
                           const float* perturbed, const float* uniform, uint64_t seed, float* child_unperturbed, float* child_perturbed,
                           void* stream);

/* ---------------------------------------------------------------------------------------------
 * Operators on HOST memory for CPU tensors (host threads, no device work).  These take host
 * pointers, allocate host memory for their indexes and return when the result is written.
 * threads = host threads to use (<= 0: all hardware threads).
 * ------------------------------------------------------------------------------------------- */
/* apply_within on host memory: psi_j in configs_j's row order, as qmb_apply_within.
 * Replaces apply_within_interface, qmb/_hamiltonian_cpu.cpp:164-257 (registered for the CPU key at
 * :558-561). */
int qmb_cpu_apply_within(int n_qubytes, int particle_cut, const uint8_t* configs_i, const double* psi_i, int64_t n_i,
                         const uint8_t* configs_j, int64_t n_j, const int16_t* site, const uint8_t* kind, const double* coef,
                         int64_t n_terms, double* psi_j, int threads);
/* The adjoint on host memory, as qmb_apply_within_adjoint (no reference counterpart: the reference
 * never differentiates the path). */
int qmb_cpu_apply_within_adjoint(int n_qubytes, int particle_cut, const uint8_t* configs_i, int64_t n_i, const uint8_t* configs_j,
                                 const double* grad_j, int64_t n_j, const int16_t* site, const uint8_t* kind, const double* coef,
                                 int64_t n_terms, double* grad_i, int threads);
/* find_relative on host memory, as qmb_find_relative; *out_rows is a HOST int64.
 * Replaces find_relative_interface, qmb/_hamiltonian_cpu.cpp:442-546 (CPU key, :558-561). */
int qmb_cpu_find_relative(int n_qubytes, int particle_cut, const uint8_t* configs_i, const double* psi_i, int64_t n_i,
                          int64_t count_selected, const int16_t* site, const uint8_t* kind, const double* coef, int64_t n_terms,
                          const uint8_t* configs_exclude, int64_t n_exclude, uint8_t* out, int64_t* out_rows, int threads);

/* Host-only pair-excitation test of a compiled record (for unit tests): returns 1 and writes
 * {r0, r1, c0, c1, delta} when the record is a pair excitation, else 0. */
int qmb_pair_excitation_host(uint64_t rec, int particle_cut, int32_t out[5]);

/* Host-only single-excitation test of a compiled record: returns 1 and writes {r, p, k, kocc, delta}
 * (k = -1 without a condition) when the record is a single excitation, else 0. */
int qmb_single_excitation_host(uint64_t rec, int particle_cut, int32_t out[5]);

#ifdef __cplusplus
}
#endif

#endif /* QMB_HIP_H */
