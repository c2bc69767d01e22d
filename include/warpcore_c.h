/*
 * warpcore_c.h -- C ABI of the MI355X-native warpcore hash-table hot path.
 *
 * The reference (sleeepyjack/warpcore) is a header-only CUDA C++ library and
 * has no C ABI; its boundary is the host half of each class template. Each
 * entry point below replaces one host member function of one explicit
 * instantiation, with the same argument meaning and the same error model:
 * nothing throws, operations are asynchronous on `stream`, per-key outcomes
 * go to the optional `status_out` (one Status word per key, bit layout of
 * include/warpcore/status.cuh:44-68) and table-wide flags accumulate in the
 * table's status word (peek_status / pop_status).
 *
 * Conventions
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream).
 *   - key/value/status/flag buffers are device pointers owned by the caller.
 *   - functions returning int return 0 on success, a hipError_t of the
 *     launch otherwise, or -1 for a NULL handle.
 *   - `probing_length` = WC_DEFAULT_PROBING_LENGTH reproduces the default
 *     argument of the reference (include/warpcore/defaults.cuh:27).
 *   - `cg_size` selects the cooperative group size (1..64 lanes of a 64-lane
 *     wavefront); 0 selects the reference's default for that class.
 */
#ifndef WARPCORE_C_H
#define WARPCORE_C_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
#define WC_API extern "C"
#else
#define WC_API
#endif

#define WC_DEFAULT_PROBING_LENGTH (~(uint64_t)0 - 1024)
#define WC_DEFAULT_SEED 0x5ad0ded

/* Status flag bits (include/warpcore/status.cuh:44-68). */
#define WC_STATUS_NONE 0u
#define WC_STATUS_UNKNOWN_ERROR (1u << 0)
#define WC_STATUS_PROBING_LENGTH_EXCEEDED (1u << 1)
#define WC_STATUS_INVALID_CONFIGURATION (1u << 2)
#define WC_STATUS_INVALID_KEY (1u << 3)
#define WC_STATUS_DUPLICATE_KEY (1u << 4)
#define WC_STATUS_KEY_NOT_FOUND (1u << 5)
#define WC_STATUS_INDEX_OVERFLOW (1u << 6)
#define WC_STATUS_OUT_OF_MEMORY (1u << 7)
#define WC_STATUS_NOT_INITIALIZED (1u << 8)
#define WC_STATUS_DRY_RUN (1u << 9)
#define WC_STATUS_INVALID_PHASE_OVERLAP (1u << 10)
#define WC_STATUS_MAX_VALUES_FOR_KEY_REACHED (1u << 11)
#define WC_STATUS_INVALID_VALUE (1u << 12)

/* ------------------------------------------------------------------------
 * SingleValueHashTable<K, V> (include/warpcore/single_value_hash_table.cuh)
 *   _create            ctor (min_capacity, seed, no_init)          :111-128
 *   _init / _init_seed init(stream) / init(seed, stream)           :175-201
 *   _insert            insert<ReturnStatus or ReturnNothing>(...)  :239-258
 *   _insert_bool       insert<ReturnBoolean<>>(...)                :239-258
 *                      (flag = has_any(~none()): invalid_value only,
 *                      status.cuh:146-150,316-333)
 *   _retrieve          retrieve<ReturnStatus or ReturnNothing>     :318-337
 *   _retrieve_bool     retrieve<ReturnBoolean<>>                   :318-337
 *   _erase             erase(...)                                  :478-496
 *   _retrieve_all      retrieve_all(keys_out, values_out, num_out) :390-417
 *   _for_each_count    for_each(f) with f adding one to *counter_d per
 *                      stored pair (device counter, caller-zeroed) :504-528
 *   _size _load_factor _capacity _bytes_total                     :534-630
 *   _peek_status _pop_status _is_initialized                       :635-687
 *   _cg_size           cg_size()                                   :100-104
 * u32u32 and u64u64 accept cg_size 1..64; u32u64 and u64u32 use 8.
 * ---------------------------------------------------------------------- */
/* SingleValueHashTable<uint32_t, uint32_t> */
WC_API void* wc_svht_u32u32_create(uint64_t min_capacity, uint32_t seed, int no_init, uint32_t cg_size);
WC_API void wc_svht_u32u32_destroy(void* table);
WC_API int wc_svht_u32u32_init(void* table, void* stream);
WC_API int wc_svht_u32u32_init_seed(void* table, uint32_t seed, void* stream);
WC_API int wc_svht_u32u32_insert(void* table, const uint32_t* keys_in, const uint32_t* values_in, uint64_t num_in, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_svht_u32u32_insert_bool(void* table, const uint32_t* keys_in, const uint32_t* values_in, uint64_t num_in, void* stream, uint64_t probing_length, bool* flags_out);
WC_API int wc_svht_u32u32_retrieve(void* table, const uint32_t* keys_in, uint64_t num_in, uint32_t* values_out, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_svht_u32u32_retrieve_bool(void* table, const uint32_t* keys_in, uint64_t num_in, uint32_t* values_out, void* stream, uint64_t probing_length, bool* flags_out);
WC_API int wc_svht_u32u32_erase(void* table, const uint32_t* keys_in, uint64_t num_in, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_svht_u32u32_retrieve_all(void* table, uint32_t* keys_out, uint32_t* values_out, uint64_t* num_out, void* stream);
WC_API int wc_svht_u32u32_for_each_count(void* table, uint64_t* counter_d, void* stream);
WC_API uint64_t wc_svht_u32u32_size(void* table, void* stream);
WC_API float wc_svht_u32u32_load_factor(void* table, void* stream);
WC_API uint64_t wc_svht_u32u32_capacity(void* table);
WC_API uint64_t wc_svht_u32u32_bytes_total(void* table);
WC_API uint32_t wc_svht_u32u32_peek_status(void* table, void* stream);
WC_API uint32_t wc_svht_u32u32_pop_status(void* table, void* stream);
WC_API int wc_svht_u32u32_is_initialized(void* table);
WC_API uint32_t wc_svht_u32u32_cg_size(void* table);

/* SingleValueHashTable<uint64_t, uint64_t> */
WC_API void* wc_svht_u64u64_create(uint64_t min_capacity, uint64_t seed, int no_init, uint32_t cg_size);
WC_API void wc_svht_u64u64_destroy(void* table);
WC_API int wc_svht_u64u64_init(void* table, void* stream);
WC_API int wc_svht_u64u64_init_seed(void* table, uint64_t seed, void* stream);
WC_API int wc_svht_u64u64_insert(void* table, const uint64_t* keys_in, const uint64_t* values_in, uint64_t num_in, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_svht_u64u64_insert_bool(void* table, const uint64_t* keys_in, const uint64_t* values_in, uint64_t num_in, void* stream, uint64_t probing_length, bool* flags_out);
WC_API int wc_svht_u64u64_retrieve(void* table, const uint64_t* keys_in, uint64_t num_in, uint64_t* values_out, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_svht_u64u64_retrieve_bool(void* table, const uint64_t* keys_in, uint64_t num_in, uint64_t* values_out, void* stream, uint64_t probing_length, bool* flags_out);
WC_API int wc_svht_u64u64_erase(void* table, const uint64_t* keys_in, uint64_t num_in, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_svht_u64u64_retrieve_all(void* table, uint64_t* keys_out, uint64_t* values_out, uint64_t* num_out, void* stream);
WC_API int wc_svht_u64u64_for_each_count(void* table, uint64_t* counter_d, void* stream);
WC_API uint64_t wc_svht_u64u64_size(void* table, void* stream);
WC_API float wc_svht_u64u64_load_factor(void* table, void* stream);
WC_API uint64_t wc_svht_u64u64_capacity(void* table);
WC_API uint64_t wc_svht_u64u64_bytes_total(void* table);
WC_API uint32_t wc_svht_u64u64_peek_status(void* table, void* stream);
WC_API uint32_t wc_svht_u64u64_pop_status(void* table, void* stream);
WC_API int wc_svht_u64u64_is_initialized(void* table);
WC_API uint32_t wc_svht_u64u64_cg_size(void* table);

/* SingleValueHashTable<uint32_t, uint64_t> */
WC_API void* wc_svht_u32u64_create(uint64_t min_capacity, uint32_t seed, int no_init, uint32_t cg_size);
WC_API void wc_svht_u32u64_destroy(void* table);
WC_API int wc_svht_u32u64_init(void* table, void* stream);
WC_API int wc_svht_u32u64_init_seed(void* table, uint32_t seed, void* stream);
WC_API int wc_svht_u32u64_insert(void* table, const uint32_t* keys_in, const uint64_t* values_in, uint64_t num_in, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_svht_u32u64_insert_bool(void* table, const uint32_t* keys_in, const uint64_t* values_in, uint64_t num_in, void* stream, uint64_t probing_length, bool* flags_out);
WC_API int wc_svht_u32u64_retrieve(void* table, const uint32_t* keys_in, uint64_t num_in, uint64_t* values_out, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_svht_u32u64_retrieve_bool(void* table, const uint32_t* keys_in, uint64_t num_in, uint64_t* values_out, void* stream, uint64_t probing_length, bool* flags_out);
WC_API int wc_svht_u32u64_erase(void* table, const uint32_t* keys_in, uint64_t num_in, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_svht_u32u64_retrieve_all(void* table, uint32_t* keys_out, uint64_t* values_out, uint64_t* num_out, void* stream);
WC_API int wc_svht_u32u64_for_each_count(void* table, uint64_t* counter_d, void* stream);
WC_API uint64_t wc_svht_u32u64_size(void* table, void* stream);
WC_API float wc_svht_u32u64_load_factor(void* table, void* stream);
WC_API uint64_t wc_svht_u32u64_capacity(void* table);
WC_API uint64_t wc_svht_u32u64_bytes_total(void* table);
WC_API uint32_t wc_svht_u32u64_peek_status(void* table, void* stream);
WC_API uint32_t wc_svht_u32u64_pop_status(void* table, void* stream);
WC_API int wc_svht_u32u64_is_initialized(void* table);
WC_API uint32_t wc_svht_u32u64_cg_size(void* table);

/* SingleValueHashTable<uint64_t, uint32_t> */
WC_API void* wc_svht_u64u32_create(uint64_t min_capacity, uint64_t seed, int no_init, uint32_t cg_size);
WC_API void wc_svht_u64u32_destroy(void* table);
WC_API int wc_svht_u64u32_init(void* table, void* stream);
WC_API int wc_svht_u64u32_init_seed(void* table, uint64_t seed, void* stream);
WC_API int wc_svht_u64u32_insert(void* table, const uint64_t* keys_in, const uint32_t* values_in, uint64_t num_in, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_svht_u64u32_insert_bool(void* table, const uint64_t* keys_in, const uint32_t* values_in, uint64_t num_in, void* stream, uint64_t probing_length, bool* flags_out);
WC_API int wc_svht_u64u32_retrieve(void* table, const uint64_t* keys_in, uint64_t num_in, uint32_t* values_out, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_svht_u64u32_retrieve_bool(void* table, const uint64_t* keys_in, uint64_t num_in, uint32_t* values_out, void* stream, uint64_t probing_length, bool* flags_out);
WC_API int wc_svht_u64u32_erase(void* table, const uint64_t* keys_in, uint64_t num_in, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_svht_u64u32_retrieve_all(void* table, uint64_t* keys_out, uint32_t* values_out, uint64_t* num_out, void* stream);
WC_API int wc_svht_u64u32_for_each_count(void* table, uint64_t* counter_d, void* stream);
WC_API uint64_t wc_svht_u64u32_size(void* table, void* stream);
WC_API float wc_svht_u64u32_load_factor(void* table, void* stream);
WC_API uint64_t wc_svht_u64u32_capacity(void* table);
WC_API uint64_t wc_svht_u64u32_bytes_total(void* table);
WC_API uint32_t wc_svht_u64u32_peek_status(void* table, void* stream);
WC_API uint32_t wc_svht_u64u32_pop_status(void* table, void* stream);
WC_API int wc_svht_u64u32_is_initialized(void* table);
WC_API uint32_t wc_svht_u64u32_cg_size(void* table);


/* ------------------------------------------------------------------------
 * SingleValueHashTable with another ProbingScheme, Hasher or TableStorage
 * template argument (include/warpcore/probing_schemes.cuh:22-302,
 * hashers.cuh:14-175, storage.cuh:218-645). The handle is that of
 * wc_svht_*_create and takes every wc_svht_* entry point of its (key,
 * value) pair. Returns NULL for a combination that is not compiled; the
 * compiled ones, all with cg_size 8 (0 = 8) and hasher1 == hasher2:
 *   probing DOUBLE    hasher MURMUR, MUELLER*, NVIDIA*, IDENTITY   storage AOS
 *   probing LINEAR    hasher MURMUR, IDENTITY                      storage AOS
 *   probing QUADRATIC hasher MURMUR, IDENTITY                      storage AOS
 *   probing DOUBLE    hasher MURMUR                                storage SOA
 * (* 32-bit keys only, as in the reference). Windows are aligned to cg
 * slots as for the default scheme; LINEAR steps one window, QUADRATIC
 * steps 1, 2, 3, ... windows.
 * ---------------------------------------------------------------------- */
#define WC_PROBING_DOUBLE 0
#define WC_PROBING_LINEAR 1
#define WC_PROBING_QUADRATIC 2
#define WC_HASHER_MURMUR 0
#define WC_HASHER_MUELLER 1
#define WC_HASHER_NVIDIA 2
#define WC_HASHER_IDENTITY 3
#define WC_STORAGE_AOS 0
#define WC_STORAGE_SOA 1
WC_API void* wc_svht_u32u32_create_custom(uint64_t min_capacity, uint32_t seed, int no_init, uint32_t cg_size, int probing, int hasher1, int hasher2, int storage);
WC_API void* wc_svht_u64u64_create_custom(uint64_t min_capacity, uint64_t seed, int no_init, uint32_t cg_size, int probing, int hasher1, int hasher2, int storage);

/* ------------------------------------------------------------------------
 * Status algebra evaluated on the host (include/warpcore/status.cuh:16-333),
 * bit for bit the reference's: all() = 0xFFF (invalid_value is outside it),
 * ~s = (~s) ^ 0xFFF, a - b = a & ~b, a + b = a | b.
 *   _return_boolean    the flag status_handlers::ReturnBoolean<ignore>
 *                      writes: has_any(~Status(ignore)) (:316-333); with
 *                      ignore = 0 it is set only by invalid_value.
 *   _return_any_flag   status_handlers::ReturnAnyFlag<ignore> (no reference
 *                      counterpart): any of the 13 flags outside ignore.
 *   _to_string         operator<< text (:216-269); returns its length.
 * ---------------------------------------------------------------------- */
WC_API uint32_t wc_status_all(void);
WC_API uint32_t wc_status_error_mask(void);
WC_API uint32_t wc_status_warning_mask(void);
WC_API uint32_t wc_status_complement(uint32_t status);
WC_API uint32_t wc_status_join(uint32_t a, uint32_t b);
WC_API uint32_t wc_status_remove(uint32_t a, uint32_t b);
WC_API int wc_status_has_any(uint32_t status, uint32_t flags);
WC_API int wc_status_has_all(uint32_t status, uint32_t flags);
WC_API int wc_status_has_any_errors(uint32_t status);
WC_API int wc_status_has_any_warnings(uint32_t status);
WC_API uint32_t wc_status_get_errors(uint32_t status);
WC_API uint32_t wc_status_get_warnings(uint32_t status);
WC_API int wc_status_return_boolean(uint32_t status, uint32_t ignore);
WC_API int wc_status_return_any_flag(uint32_t status, uint32_t ignore);
WC_API int wc_status_to_string(uint32_t status, char* buf, uint64_t len);

/* ------------------------------------------------------------------------
 * CountingHashTable<K, uint64_t> (include/warpcore/counting_hash_table.cuh)
 *   _create (min_capacity, seed, max_count, no_init)               :74-85
 *   _init / _init_seed                                              :113-135
 *   _insert   insert(keys, n, stream, probing_length, status_out)  :182-194
 *   _retrieve counts_out[i] = min(count, max_count) on a hit       :236-249
 *   _erase    erase(keys, n, stream, probing_length, status_out)   :273-301
 *   _retrieve_all / _size / _load_factor / _capacity / _peek_status /
 *   _pop_status                                                    :262-390
 * cg_size 1..64, default 4.
 * ---------------------------------------------------------------------- */
/* CountingHashTable<uint32_t, uint64_t> */
WC_API void* wc_counting_u32_create(uint64_t min_capacity, uint32_t seed, uint64_t max_count, int no_init, uint32_t cg_size);
WC_API void wc_counting_u32_destroy(void* table);
WC_API int wc_counting_u32_init(void* table, void* stream);
WC_API int wc_counting_u32_init_seed(void* table, uint32_t seed, void* stream);
WC_API int wc_counting_u32_insert(void* table, const uint32_t* keys_in, uint64_t num_in, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_counting_u32_retrieve(void* table, const uint32_t* keys_in, uint64_t num_in, uint64_t* counts_out, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_counting_u32_erase(void* table, const uint32_t* keys_in, uint64_t num_in, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_counting_u32_retrieve_all(void* table, uint32_t* keys_out, uint64_t* counts_out, uint64_t* num_out, void* stream);
WC_API uint64_t wc_counting_u32_size(void* table, void* stream);
WC_API float wc_counting_u32_load_factor(void* table, void* stream);
WC_API uint64_t wc_counting_u32_capacity(void* table);
WC_API uint64_t wc_counting_u32_bytes_total(void* table);
WC_API uint32_t wc_counting_u32_peek_status(void* table, void* stream);
WC_API uint32_t wc_counting_u32_pop_status(void* table, void* stream);
WC_API uint32_t wc_counting_u32_cg_size(void* table);

/* CountingHashTable<uint64_t, uint64_t> */
WC_API void* wc_counting_u64_create(uint64_t min_capacity, uint64_t seed, uint64_t max_count, int no_init, uint32_t cg_size);
WC_API void wc_counting_u64_destroy(void* table);
WC_API int wc_counting_u64_init(void* table, void* stream);
WC_API int wc_counting_u64_init_seed(void* table, uint64_t seed, void* stream);
WC_API int wc_counting_u64_insert(void* table, const uint64_t* keys_in, uint64_t num_in, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_counting_u64_retrieve(void* table, const uint64_t* keys_in, uint64_t num_in, uint64_t* counts_out, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_counting_u64_erase(void* table, const uint64_t* keys_in, uint64_t num_in, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_counting_u64_retrieve_all(void* table, uint64_t* keys_out, uint64_t* counts_out, uint64_t* num_out, void* stream);
WC_API uint64_t wc_counting_u64_size(void* table, void* stream);
WC_API float wc_counting_u64_load_factor(void* table, void* stream);
WC_API uint64_t wc_counting_u64_capacity(void* table);
WC_API uint64_t wc_counting_u64_bytes_total(void* table);
WC_API uint32_t wc_counting_u64_peek_status(void* table, void* stream);
WC_API uint32_t wc_counting_u64_pop_status(void* table, void* stream);
WC_API uint32_t wc_counting_u64_cg_size(void* table);

/*
 * wc_counting_u64_insert_noalloc -- no counterpart in the reference. For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc): the same results,
 * counts, per-occurrence statuses (which occurrence of a new key reports
 * `none` is unspecified, as in _insert) and table status as
 * wc_counting_u64_insert, with the combining buffers in a fixed device array
 * of the library (30 MiB per GPU, held until the process ends) instead of
 * memory the table allocates, and no host synchronisation. It calls no
 * hipMalloc/hipFree and no allocating library routine.
 *   num_in at most wc_counting_u64_insert_noalloc_max() (2^20); larger
 *   inputs return WC_ERR_TOO_LARGE (-4) and change nothing.
 *   Returns WC_ERR_NO_SCRATCH (-5) when the array's device address cannot
 *   be obtained, else 0 or the hipError_t of the launch, -1 for NULL.
 *   Calls that overlap on one GPU (other streams, other host threads) share
 *   the array and race; calls one after another on one stream do not.
 */
#define WC_ERR_TOO_LARGE (-4)
#define WC_ERR_NO_SCRATCH (-5)
WC_API int wc_counting_u64_insert_noalloc(void* table, const uint64_t* keys_in, uint64_t num_in, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API uint64_t wc_counting_u64_insert_noalloc_max(void);

/* ------------------------------------------------------------------------
 * HashSet<K> (include/warpcore/hash_set.cuh)
 *   _create (min_capacity, seed)                                   :80-110
 *   _insert                                                        :264-282
 *   _contains flags_out[i] = membership, written when the status is none;
 *             the device retrieve semantics of :291-329 (the host retrieve
 *             of :340-359 cannot be instantiated)
 *   _erase (status none whatever it finds)                         :446-462
 *   _for_each_count for_each(f), f adding one to *counter_d per key  :472
 *   _retrieve_all / _size / _load_factor / _capacity / _peek_status /
 *   _pop_status                                                    :366-640
 * cg_size 1..64, default 16.
 * ---------------------------------------------------------------------- */
/* HashSet<uint32_t> */
WC_API void* wc_hashset_u32_create(uint64_t min_capacity, uint32_t seed, uint32_t cg_size);
WC_API void wc_hashset_u32_destroy(void* set);
WC_API int wc_hashset_u32_init(void* set, void* stream);
WC_API int wc_hashset_u32_init_seed(void* set, uint32_t seed, void* stream);
WC_API int wc_hashset_u32_insert(void* set, const uint32_t* keys_in, uint64_t num_in, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_hashset_u32_contains(void* set, const uint32_t* keys_in, uint64_t num_in, bool* flags_out, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_hashset_u32_erase(void* set, const uint32_t* keys_in, uint64_t num_in, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_hashset_u32_retrieve_all(void* set, uint32_t* keys_out, uint64_t* num_out, void* stream);
WC_API int wc_hashset_u32_for_each_count(void* set, uint64_t* counter_d, void* stream);
WC_API uint64_t wc_hashset_u32_size(void* set, void* stream);
WC_API float wc_hashset_u32_load_factor(void* set, void* stream);
WC_API uint64_t wc_hashset_u32_capacity(void* set);
WC_API uint64_t wc_hashset_u32_bytes_total(void* set);
WC_API uint32_t wc_hashset_u32_peek_status(void* set, void* stream);
WC_API uint32_t wc_hashset_u32_pop_status(void* set, void* stream);
WC_API uint32_t wc_hashset_u32_cg_size(void* set);

/* HashSet<uint64_t> */
WC_API void* wc_hashset_u64_create(uint64_t min_capacity, uint64_t seed, uint32_t cg_size);
WC_API void wc_hashset_u64_destroy(void* set);
WC_API int wc_hashset_u64_init(void* set, void* stream);
WC_API int wc_hashset_u64_init_seed(void* set, uint64_t seed, void* stream);
WC_API int wc_hashset_u64_insert(void* set, const uint64_t* keys_in, uint64_t num_in, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_hashset_u64_contains(void* set, const uint64_t* keys_in, uint64_t num_in, bool* flags_out, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_hashset_u64_erase(void* set, const uint64_t* keys_in, uint64_t num_in, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_hashset_u64_retrieve_all(void* set, uint64_t* keys_out, uint64_t* num_out, void* stream);
WC_API int wc_hashset_u64_for_each_count(void* set, uint64_t* counter_d, void* stream);
WC_API uint64_t wc_hashset_u64_size(void* set, void* stream);
WC_API float wc_hashset_u64_load_factor(void* set, void* stream);
WC_API uint64_t wc_hashset_u64_capacity(void* set);
WC_API uint64_t wc_hashset_u64_bytes_total(void* set);
WC_API uint32_t wc_hashset_u64_peek_status(void* set, void* stream);
WC_API uint32_t wc_hashset_u64_pop_status(void* set, void* stream);
WC_API uint32_t wc_hashset_u64_cg_size(void* set);

/* ------------------------------------------------------------------------
 * BloomFilter<K, MurmurHash<K>, Slot, CG> (include/warpcore/bloom_filter.cuh)
 *   create       constructor :77-93; slot_bits 32 or 64 (0 = 64), cg_size
 *                1..64 power of two (0 = 1); NULL for num_bits == 0, k == 0
 *                or an unsupported configuration
 *   init         :142-161 (clears all bits; _init_seed also re-seeds)
 *   insert       bulk insert :199-208 (device member :167-192)
 *   retrieve     bulk query :245-255 (device member :215-237); flags_out[i]
 *                is 1 iff every bit of keys_in[i] is set
 *   insert_and_query  bulk form of the device member :263-297, cg_size 1
 *                only (else -2); flag = all bits were set before the key's
 *                own update
 *   copy_words   copies the num_slots words of the bit array to a device
 *                buffer of bytes_total bytes (no reference counterpart; it
 *                exposes the array the reference keeps private)
 *   fpr          blocked-model false-positive estimate :341-408, summed in
 *                log space so it stays finite for large n
 *   num_bits/num_slots/num_blocks/k :304-338
 * ---------------------------------------------------------------------- */
/* BloomFilter<uint32_t> */
WC_API void* wc_bloom_u32_create(uint64_t num_bits, uint64_t k, uint32_t seed, uint32_t slot_bits, uint32_t cg_size);
WC_API void wc_bloom_u32_destroy(void* filter);
WC_API int wc_bloom_u32_init(void* filter, void* stream);
WC_API int wc_bloom_u32_init_seed(void* filter, uint32_t seed, void* stream);
WC_API int wc_bloom_u32_insert(void* filter, const uint32_t* keys_in, uint64_t num_in, void* stream);
WC_API int wc_bloom_u32_retrieve(void* filter, const uint32_t* keys_in, uint64_t num_in, bool* flags_out, void* stream);
WC_API int wc_bloom_u32_insert_and_query(void* filter, const uint32_t* keys_in, uint64_t num_in, bool* flags_out, void* stream);
WC_API int wc_bloom_u32_copy_words(void* filter, void* words_out, void* stream);
WC_API uint64_t wc_bloom_u32_num_bits(void* filter);
WC_API uint64_t wc_bloom_u32_num_slots(void* filter);
WC_API uint64_t wc_bloom_u32_num_blocks(void* filter);
WC_API uint64_t wc_bloom_u32_k(void* filter);
WC_API uint32_t wc_bloom_u32_seed(void* filter);
WC_API uint32_t wc_bloom_u32_slot_bits(void* filter);
WC_API uint32_t wc_bloom_u32_cg_size(void* filter);
WC_API uint64_t wc_bloom_u32_bytes_total(void* filter);
WC_API double wc_bloom_u32_fpr(void* filter, uint64_t n);

/* BloomFilter<uint64_t> */
WC_API void* wc_bloom_u64_create(uint64_t num_bits, uint64_t k, uint64_t seed, uint32_t slot_bits, uint32_t cg_size);
WC_API void wc_bloom_u64_destroy(void* filter);
WC_API int wc_bloom_u64_init(void* filter, void* stream);
WC_API int wc_bloom_u64_init_seed(void* filter, uint64_t seed, void* stream);
WC_API int wc_bloom_u64_insert(void* filter, const uint64_t* keys_in, uint64_t num_in, void* stream);
WC_API int wc_bloom_u64_retrieve(void* filter, const uint64_t* keys_in, uint64_t num_in, bool* flags_out, void* stream);
WC_API int wc_bloom_u64_insert_and_query(void* filter, const uint64_t* keys_in, uint64_t num_in, bool* flags_out, void* stream);
WC_API int wc_bloom_u64_copy_words(void* filter, void* words_out, void* stream);
WC_API uint64_t wc_bloom_u64_num_bits(void* filter);
WC_API uint64_t wc_bloom_u64_num_slots(void* filter);
WC_API uint64_t wc_bloom_u64_num_blocks(void* filter);
WC_API uint64_t wc_bloom_u64_k(void* filter);
WC_API uint64_t wc_bloom_u64_seed(void* filter);
WC_API uint32_t wc_bloom_u64_slot_bits(void* filter);
WC_API uint32_t wc_bloom_u64_cg_size(void* filter);
WC_API uint64_t wc_bloom_u64_bytes_total(void* filter);
WC_API double wc_bloom_u64_fpr(void* filter, uint64_t n);

/* ------------------------------------------------------------------------
 * MultiValueHashTable<K, V> (include/warpcore/multi_value_hash_table.cuh)
 *   create       constructor :93-123 (max_values_per_key, no_init); cg_size
 *                1..64 power of two (0 = 8, the reference default)
 *   init         :160-177 (empties the table, zeroes the key counter)
 *   insert       :297-319 (device member :200-289); status per pair: none for
 *                a key's first value, duplicate_key for later ones,
 *                duplicate_key|max_values_for_key_reached when refused,
 *                probing_length_exceeded, invalid_key
 *   retrieve     two-pass retrieve :370-510: end_offsets_out[i] = inclusive
 *                prefix sum of the counts, begin_offsets_out[i] = its
 *                exclusive sum, values of keys_in[i] at
 *                values_out[begin[i] .. end[i]) in probe order, *num_out =
 *                total; values_out == NULL is a dry run: end_offsets_out
 *                gets the raw counts, status dry_run
 *   num_values   :766-793 per-key counts (num_per_key_out may be NULL) and
 *                their sum
 *   retrieve_all_keys :562-600 each distinct key once (order unspecified)
 *   retrieve_all :524-548 retrieve_all_keys, then retrieve for those keys
 *   num_keys :721-731, size = num_values :809-834, load_factor,
 *   storage_density :851-862, peek_status/pop_status :884-925
 * ---------------------------------------------------------------------- */
/* MultiValueHashTable<uint32_t, uint32_t> */
WC_API void* wc_mvht_u32u32_create(uint64_t min_capacity, uint32_t seed, uint64_t max_values_per_key, int no_init, uint32_t cg_size);
WC_API void wc_mvht_u32u32_destroy(void* table);
WC_API int wc_mvht_u32u32_init(void* table, void* stream);
WC_API int wc_mvht_u32u32_insert(void* table, const uint32_t* keys_in, const uint32_t* values_in, uint64_t num_in, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_mvht_u32u32_retrieve(void* table, const uint32_t* keys_in, uint64_t num_in, uint64_t* begin_offsets_out, uint64_t* end_offsets_out, uint32_t* values_out, uint64_t* num_out, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_mvht_u32u32_num_values(void* table, const uint32_t* keys_in, uint64_t num_in, uint64_t* num_out, uint64_t* num_per_key_out, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_mvht_u32u32_retrieve_all_keys(void* table, uint32_t* keys_out, uint64_t* num_out, void* stream);
WC_API int wc_mvht_u32u32_retrieve_all(void* table, uint32_t* keys_out, uint64_t* num_keys_out, uint64_t* begin_offsets_out, uint64_t* end_offsets_out, uint32_t* values_out, uint64_t* num_values_out, void* stream);
WC_API uint64_t wc_mvht_u32u32_num_keys(void* table, void* stream);
WC_API uint64_t wc_mvht_u32u32_size(void* table, void* stream);
WC_API float wc_mvht_u32u32_load_factor(void* table, void* stream);
WC_API float wc_mvht_u32u32_storage_density(void* table, void* stream);
WC_API uint64_t wc_mvht_u32u32_capacity(void* table);
WC_API uint64_t wc_mvht_u32u32_bytes_total(void* table);
WC_API uint64_t wc_mvht_u32u32_max_values_per_key(void* table);
WC_API int wc_mvht_u32u32_is_initialized(void* table);
WC_API uint32_t wc_mvht_u32u32_peek_status(void* table, void* stream);
WC_API uint32_t wc_mvht_u32u32_pop_status(void* table, void* stream);
WC_API uint32_t wc_mvht_u32u32_cg_size(void* table);

/* MultiValueHashTable<uint64_t, uint64_t> */
WC_API void* wc_mvht_u64u64_create(uint64_t min_capacity, uint64_t seed, uint64_t max_values_per_key, int no_init, uint32_t cg_size);
WC_API void wc_mvht_u64u64_destroy(void* table);
WC_API int wc_mvht_u64u64_init(void* table, void* stream);
WC_API int wc_mvht_u64u64_insert(void* table, const uint64_t* keys_in, const uint64_t* values_in, uint64_t num_in, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_mvht_u64u64_retrieve(void* table, const uint64_t* keys_in, uint64_t num_in, uint64_t* begin_offsets_out, uint64_t* end_offsets_out, uint64_t* values_out, uint64_t* num_out, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_mvht_u64u64_num_values(void* table, const uint64_t* keys_in, uint64_t num_in, uint64_t* num_out, uint64_t* num_per_key_out, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_mvht_u64u64_retrieve_all_keys(void* table, uint64_t* keys_out, uint64_t* num_out, void* stream);
WC_API int wc_mvht_u64u64_retrieve_all(void* table, uint64_t* keys_out, uint64_t* num_keys_out, uint64_t* begin_offsets_out, uint64_t* end_offsets_out, uint64_t* values_out, uint64_t* num_values_out, void* stream);
WC_API uint64_t wc_mvht_u64u64_num_keys(void* table, void* stream);
WC_API uint64_t wc_mvht_u64u64_size(void* table, void* stream);
WC_API float wc_mvht_u64u64_load_factor(void* table, void* stream);
WC_API float wc_mvht_u64u64_storage_density(void* table, void* stream);
WC_API uint64_t wc_mvht_u64u64_capacity(void* table);
WC_API uint64_t wc_mvht_u64u64_bytes_total(void* table);
WC_API uint64_t wc_mvht_u64u64_max_values_per_key(void* table);
WC_API int wc_mvht_u64u64_is_initialized(void* table);
WC_API uint32_t wc_mvht_u64u64_peek_status(void* table, void* stream);
WC_API uint32_t wc_mvht_u64u64_pop_status(void* table, void* stream);
WC_API uint32_t wc_mvht_u64u64_cg_size(void* table);

/* ------------------------------------------------------------------------
 * BucketListHashTable<K, V> (include/warpcore/bucket_list_hash_table.cuh)
 * with storage::multi_value::BucketListStore<V> (storage.cuh:656-1454)
 *   create       constructor :106-125: key_capacity (minimum, capacity rule
 *                applied), value_capacity (pool cells), grow_factor >= 1,
 *                1 <= min_bucket_size <= max_bucket_size <= 1023,
 *                max_values_per_key (clamped to 2^20 - 1); cg_size 0 = 8.
 *                An invalid pool configuration leaves the table
 *                uninitialised (status invalid_configuration|not_initialized)
 *   init         :155-171 (also empties the value pool)
 *   insert       status per pair = key-table status (none for a new key,
 *                duplicate_key for a known one) + append status
 *                (out_of_memory, max_values_for_key_reached)
 *   retrieve / num_values / retrieve_all_keys / retrieve_all: as for
 *                MultiValueHashTable; values of a key come tail bucket first
 *   total_values :805-822 (num_values() of the whole table)
 *   key_load_factor / value_load_factor / storage_density /
 *   relative_storage_density / key_capacity / value_capacity :612-737
 *   value_cells_used  pool cells handed out so far (no reference
 *                counterpart; value_load_factor * value_capacity)
 * ---------------------------------------------------------------------- */
/* BucketListHashTable<uint32_t, uint32_t> */
WC_API void* wc_bucketlist_u32u32_create(uint64_t key_capacity, uint64_t value_capacity, uint32_t seed, float grow_factor, uint64_t min_bucket_size, uint64_t max_bucket_size, uint64_t max_values_per_key, int no_init, uint32_t cg_size);
WC_API void wc_bucketlist_u32u32_destroy(void* table);
WC_API int wc_bucketlist_u32u32_init(void* table, void* stream);
WC_API int wc_bucketlist_u32u32_init_seed(void* table, uint32_t seed, void* stream);
WC_API int wc_bucketlist_u32u32_insert(void* table, const uint32_t* keys_in, const uint32_t* values_in, uint64_t num_in, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_bucketlist_u32u32_retrieve(void* table, const uint32_t* keys_in, uint64_t num_in, uint64_t* begin_offsets_out, uint64_t* end_offsets_out, uint32_t* values_out, uint64_t* num_out, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_bucketlist_u32u32_num_values(void* table, const uint32_t* keys_in, uint64_t num_in, uint64_t* num_out, uint64_t* num_per_key_out, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API uint64_t wc_bucketlist_u32u32_total_values(void* table, void* stream);
WC_API int wc_bucketlist_u32u32_retrieve_all_keys(void* table, uint32_t* keys_out, uint64_t* num_out, void* stream);
WC_API int wc_bucketlist_u32u32_retrieve_all(void* table, uint32_t* keys_out, uint64_t* num_keys_out, uint64_t* begin_offsets_out, uint64_t* end_offsets_out, uint32_t* values_out, uint64_t* num_values_out, void* stream);
WC_API uint64_t wc_bucketlist_u32u32_num_keys(void* table, void* stream);
WC_API float wc_bucketlist_u32u32_key_load_factor(void* table, void* stream);
WC_API float wc_bucketlist_u32u32_value_load_factor(void* table, void* stream);
WC_API float wc_bucketlist_u32u32_storage_density(void* table, void* stream);
WC_API float wc_bucketlist_u32u32_relative_storage_density(void* table, void* stream);
WC_API uint64_t wc_bucketlist_u32u32_key_capacity(void* table);
WC_API uint64_t wc_bucketlist_u32u32_value_capacity(void* table);
WC_API uint64_t wc_bucketlist_u32u32_value_cells_used(void* table, void* stream);
WC_API uint64_t wc_bucketlist_u32u32_bytes_total(void* table);
WC_API uint64_t wc_bucketlist_u32u32_max_values_per_key(void* table);
WC_API int wc_bucketlist_u32u32_is_initialized(void* table);
WC_API uint32_t wc_bucketlist_u32u32_peek_status(void* table, void* stream);
WC_API uint32_t wc_bucketlist_u32u32_pop_status(void* table, void* stream);
WC_API uint32_t wc_bucketlist_u32u32_cg_size(void* table);

/* BucketListHashTable<uint64_t, uint64_t> */
WC_API void* wc_bucketlist_u64u64_create(uint64_t key_capacity, uint64_t value_capacity, uint64_t seed, float grow_factor, uint64_t min_bucket_size, uint64_t max_bucket_size, uint64_t max_values_per_key, int no_init, uint32_t cg_size);
WC_API void wc_bucketlist_u64u64_destroy(void* table);
WC_API int wc_bucketlist_u64u64_init(void* table, void* stream);
WC_API int wc_bucketlist_u64u64_init_seed(void* table, uint64_t seed, void* stream);
WC_API int wc_bucketlist_u64u64_insert(void* table, const uint64_t* keys_in, const uint64_t* values_in, uint64_t num_in, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_bucketlist_u64u64_retrieve(void* table, const uint64_t* keys_in, uint64_t num_in, uint64_t* begin_offsets_out, uint64_t* end_offsets_out, uint64_t* values_out, uint64_t* num_out, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_bucketlist_u64u64_num_values(void* table, const uint64_t* keys_in, uint64_t num_in, uint64_t* num_out, uint64_t* num_per_key_out, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API uint64_t wc_bucketlist_u64u64_total_values(void* table, void* stream);
WC_API int wc_bucketlist_u64u64_retrieve_all_keys(void* table, uint64_t* keys_out, uint64_t* num_out, void* stream);
WC_API int wc_bucketlist_u64u64_retrieve_all(void* table, uint64_t* keys_out, uint64_t* num_keys_out, uint64_t* begin_offsets_out, uint64_t* end_offsets_out, uint64_t* values_out, uint64_t* num_values_out, void* stream);
WC_API uint64_t wc_bucketlist_u64u64_num_keys(void* table, void* stream);
WC_API float wc_bucketlist_u64u64_key_load_factor(void* table, void* stream);
WC_API float wc_bucketlist_u64u64_value_load_factor(void* table, void* stream);
WC_API float wc_bucketlist_u64u64_storage_density(void* table, void* stream);
WC_API float wc_bucketlist_u64u64_relative_storage_density(void* table, void* stream);
WC_API uint64_t wc_bucketlist_u64u64_key_capacity(void* table);
WC_API uint64_t wc_bucketlist_u64u64_value_capacity(void* table);
WC_API uint64_t wc_bucketlist_u64u64_value_cells_used(void* table, void* stream);
WC_API uint64_t wc_bucketlist_u64u64_bytes_total(void* table);
WC_API uint64_t wc_bucketlist_u64u64_max_values_per_key(void* table);
WC_API int wc_bucketlist_u64u64_is_initialized(void* table);
WC_API uint32_t wc_bucketlist_u64u64_peek_status(void* table, void* stream);
WC_API uint32_t wc_bucketlist_u64u64_pop_status(void* table, void* stream);
WC_API uint32_t wc_bucketlist_u64u64_cg_size(void* table);

/* ------------------------------------------------------------------------
 * TableBatch<K, V>: many small SingleValueHashTables in one arena (no
 * reference counterpart; the reference builds one table per object with at
 * least 131101*cg slots). Each table has a configuration (lanes, window):
 * groups of `lanes` lanes probe windows of `window` slots; table t gets the
 * smallest prime >= max(2, SDIV(min, window)) windows of `window` slots.
 * Configurations, by index: u32u32 (4,4) (4,8) (8,8) (8,16) (16,16);
 * u64u64 (4,4) (8,8) (16,16).
 *   create       min_capacities is a HOST array of num_tables slot counts;
 *                every table starts at configuration (cg_size, cg_size)
 *                (cg_size 4, 8 or 16; 0 = 16)
 *   insert / retrieve  keys of all tables concatenated; offsets is a DEVICE
 *                array of num_tables + 1 (table t owns [offsets[t],
 *                offsets[t+1])); one persistent work-stealing launch; per
 *                key results and statuses as SingleValueHashTable's
 *   set_launch / get_launch  keys per work item, blocks per CU, threads per
 *                block (64..512, a multiple of 64)
 *   set_lds_insert  inserts and lookups in the tables of at most 128 KiB of
 *                slots run in LDS, one block per table, beside the work items
 *                (1, the default) or through the work items like the others
 *                (0); same results and statuses
 *   lds_tables   how many tables the LDS path takes
 *   tune         autotuner: tables are put in classes by requested capacity
 *                (bands [4^k, 4^(k+1))) and fill (keys / requested capacity:
 *                < 0.6, < 0.7, < 0.8, >= 0.8); every configuration is timed
 *                on every class alone (insert + retrieve of the given
 *                workload restricted to the class), each class keeps its
 *                fastest configuration within 10% of its fewest slots, then
 *                the launch candidates are timed on the whole workload.
 *                Leaves the batch empty with the chosen configurations and
 *                launch. launch_out: 3 values as get_launch; ms_out: tuned
 *                insert + retrieve time of the whole workload
 *   tune_report  after a tune: rows of 10 values per class (class id =
 *                capacity band * 4 + fill band, tables, keys, queries,
 *                fixed configuration, chosen configuration, fixed ns, chosen
 *                ns, fixed slots, chosen slots), at most max_rows written;
 *                totals[2] = whole workload ms at the creation configuration
 *                with the default launch, and tuned. Returns the number of
 *                classes (0 before a tune)
 *   set_configs  per-table configuration indices (HOST array of num_tables
 *                bytes); lays the tables out again and empties the batch
 *   table_config / num_configs / config_shape  a table's configuration
 *                index; the number of configurations; (lanes, window) of one
 *   cg_size      lanes of the creation configuration
 *   copy_status_words  per-table status words into a device buffer
 * ---------------------------------------------------------------------- */
/* TableBatch<uint32_t, uint32_t> */
WC_API void* wc_batch_u32u32_create(const uint64_t* min_capacities, uint64_t num_tables, uint32_t seed, uint32_t cg_size);
WC_API void wc_batch_u32u32_destroy(void* batch);
WC_API int wc_batch_u32u32_init(void* batch, void* stream);
WC_API int wc_batch_u32u32_insert(void* batch, const uint32_t* keys_in, const uint32_t* values_in, const uint64_t* offsets, uint64_t num_in, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_batch_u32u32_retrieve(void* batch, const uint32_t* keys_in, const uint64_t* offsets, uint64_t num_in, uint32_t* values_out, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_batch_u32u32_set_launch(void* batch, uint64_t item_keys, uint32_t blocks_per_cu, uint32_t block_size);
WC_API int wc_batch_u32u32_set_lds_insert(void* batch, int on);
WC_API uint64_t wc_batch_u32u32_lds_tables(void* batch);
WC_API int wc_batch_u32u32_get_launch(void* batch, uint64_t* launch_out);
WC_API int wc_batch_u32u32_tune(void* batch, const uint32_t* keys_in, const uint32_t* values_in, const uint64_t* offsets, uint64_t num_in, const uint32_t* queries, const uint64_t* query_offsets, uint64_t num_queries, uint32_t* values_out, void* stream, uint64_t* launch_out, double* ms_out);
WC_API int wc_batch_u32u32_tune_report(void* batch, uint64_t* rows, uint64_t max_rows, double* totals);
WC_API int wc_batch_u32u32_set_configs(void* batch, const uint8_t* configs, uint64_t num_tables);
WC_API int wc_batch_u32u32_table_config(void* batch, uint64_t table);
WC_API int wc_batch_u32u32_num_configs(void);
WC_API int wc_batch_u32u32_config_shape(uint32_t config, uint32_t* lanes_window);
WC_API uint64_t wc_batch_u32u32_num_tables(void* batch);
WC_API uint64_t wc_batch_u32u32_capacity(void* batch, uint64_t table);
WC_API uint64_t wc_batch_u32u32_total_slots(void* batch);
WC_API uint64_t wc_batch_u32u32_bytes_total(void* batch);
WC_API int wc_batch_u32u32_copy_status_words(void* batch, uint32_t* words_out, void* stream);
WC_API uint32_t wc_batch_u32u32_cg_size(void* batch);

/* TableBatch<uint64_t, uint64_t> */
WC_API void* wc_batch_u64u64_create(const uint64_t* min_capacities, uint64_t num_tables, uint64_t seed, uint32_t cg_size);
WC_API void wc_batch_u64u64_destroy(void* batch);
WC_API int wc_batch_u64u64_init(void* batch, void* stream);
WC_API int wc_batch_u64u64_insert(void* batch, const uint64_t* keys_in, const uint64_t* values_in, const uint64_t* offsets, uint64_t num_in, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_batch_u64u64_retrieve(void* batch, const uint64_t* keys_in, const uint64_t* offsets, uint64_t num_in, uint64_t* values_out, void* stream, uint64_t probing_length, uint32_t* status_out);
WC_API int wc_batch_u64u64_set_launch(void* batch, uint64_t item_keys, uint32_t blocks_per_cu, uint32_t block_size);
WC_API int wc_batch_u64u64_set_lds_insert(void* batch, int on);
WC_API uint64_t wc_batch_u64u64_lds_tables(void* batch);
WC_API int wc_batch_u64u64_get_launch(void* batch, uint64_t* launch_out);
WC_API int wc_batch_u64u64_tune(void* batch, const uint64_t* keys_in, const uint64_t* values_in, const uint64_t* offsets, uint64_t num_in, const uint64_t* queries, const uint64_t* query_offsets, uint64_t num_queries, uint64_t* values_out, void* stream, uint64_t* launch_out, double* ms_out);
WC_API int wc_batch_u64u64_tune_report(void* batch, uint64_t* rows, uint64_t max_rows, double* totals);
WC_API int wc_batch_u64u64_set_configs(void* batch, const uint8_t* configs, uint64_t num_tables);
WC_API int wc_batch_u64u64_table_config(void* batch, uint64_t table);
WC_API int wc_batch_u64u64_num_configs(void);
WC_API int wc_batch_u64u64_config_shape(uint32_t config, uint32_t* lanes_window);
WC_API uint64_t wc_batch_u64u64_num_tables(void* batch);
WC_API uint64_t wc_batch_u64u64_capacity(void* batch, uint64_t table);
WC_API uint64_t wc_batch_u64u64_total_slots(void* batch);
WC_API uint64_t wc_batch_u64u64_bytes_total(void* batch);
WC_API int wc_batch_u64u64_copy_status_words(void* batch, uint32_t* words_out, void* stream);
WC_API uint32_t wc_batch_u64u64_cg_size(void* batch);

/* ------------------------------------------------------------------------
 * Serving loop (no reference counterpart): device-resident
 * SingleValueHashTables (u64u64: cg 8, u32u32: cg 16) serve insert (op 0),
 * retrieve (op 1) and erase (op 2) requests from many host threads. A
 * scheduler thread coalesces admitted requests into steps: one persistent
 * launch over all requests of a step, captured as a HIP graph when
 * use_graph. A step holds at most max_step_keys keys; a larger request
 * runs in chunks over consecutive steps. Per table, requests run in
 * arrival order (a step holds a run of lookups or one insert/erase, or one
 * chunk of it, per table). While fewer than 2^24 keys are pending (latency
 * mode), a request of at least direct_keys keys runs whole as a step of its
 * own on its table's bulk kernels (the kernels of one plain call).
 * Admission: each table has a budget of max_load * capacity claimed slots
 * (stored keys and tombstones). An insert of n keys reserves n slots when
 * admitted; the slots its steps did not claim (keys already stored) are
 * given back when they have run. An insert that does not fit waits for the
 * steps in flight; if it still does not fit, a table holding tombstones is
 * compacted (rebuilt from its live pairs; results do not change), and an
 * insert that does not fit even then is refused: every key reports
 * out_of_memory and wait returns 0.
 *   create       table_capacities: HOST array (minimum capacities);
 *                max_requests per step (0 = 1024), item_keys: largest work
 *                item of the step kernel (0 = 2048; a step uses the smallest
 *                power of two from 64 that gives every wavefront about four
 *                items), max_step_keys (0 = 2^24), max_load (<= 0: 0.9),
 *                blocks_per_cu (0 = 8), direct_keys (0 = 2^18, UINT64_MAX =
 *                every step on the step kernel)
 *   submit       device buffers, kept valid until the request is done;
 *                returns the request id (-1 on a bad argument)
 *   wait         block until request id is done: 1 served, 0 refused,
 *                -6 a step of the request failed to launch or complete
 *                (its outputs are undefined), -7 unknown id (never issued,
 *                or among the completions older than the last 2^20, whose
 *                outcomes are dropped); -1 NULL server
 *   drain        block until every request is done
 *   latency_us   arrival to completion of a done request (-1 otherwise)
 *   step_of      the step a request ran in (1, 2, ...)
 *   stats        out3 = {requests done, steps, keys}
 *   direct_steps steps run on a table's bulk kernels
 *   counters     out4 = {direct steps, compactions, failed requests, 0}
 *   table_reserved / table_budget  claimed + reserved slots / the budget
 *   fail_step    failure handling from the host (for tests): step number
 *                `step` (1, 2, ... in issue order) is not launched and
 *                reports a launch failure; its requests fail (-6)
 *   drive        replays num_requests requests (keys/values/outputs are
 *                slices of concatenated device arrays, sizes[i] each) at
 *                arrival_us[i] after the call starts; mode 0 through the
 *                server, mode 1 one plain table call per request in arrival
 *                order on the server's stream; fills latency_us_out[i] and
 *                *makespan_us
 * ---------------------------------------------------------------------- */
/* Server over SingleValueHashTable<uint64_t, uint64_t> */
WC_API void* wc_server_u64u64_create(const uint64_t* table_capacities, uint64_t num_tables, uint64_t max_requests, uint64_t item_keys, uint64_t max_step_keys, float max_load, int use_graph, int blocks_per_cu, uint64_t direct_keys);
WC_API void wc_server_u64u64_destroy(void* server);
WC_API int64_t wc_server_u64u64_submit(void* server, uint32_t op, uint32_t table, const uint64_t* keys_in, const uint64_t* values_in, uint64_t* values_out, uint32_t* status_out, uint64_t num_in);
WC_API int wc_server_u64u64_wait(void* server, int64_t id);
WC_API int wc_server_u64u64_drain(void* server);
WC_API double wc_server_u64u64_latency_us(void* server, int64_t id);
WC_API uint64_t wc_server_u64u64_step_of(void* server, int64_t id);
WC_API int wc_server_u64u64_stats(void* server, uint64_t* out3);
WC_API uint64_t wc_server_u64u64_direct_steps(void* server);
WC_API int wc_server_u64u64_counters(void* server, uint64_t* out4);
WC_API uint64_t wc_server_u64u64_table_reserved(void* server, uint32_t table);
WC_API uint64_t wc_server_u64u64_table_budget(void* server, uint32_t table);
WC_API int wc_server_u64u64_fail_step(void* server, uint64_t step);
WC_API uint64_t wc_server_u64u64_table_size(void* server, uint32_t table);
WC_API uint64_t wc_server_u64u64_table_capacity(void* server, uint32_t table);
WC_API uint32_t wc_server_u64u64_table_status(void* server, uint32_t table);
WC_API int wc_server_u64u64_drive(void* server, int mode, const uint32_t* ops, const uint32_t* tables, const uint64_t* sizes, const double* arrival_us, uint64_t num_requests, const uint64_t* keys_in, const uint64_t* values_in, uint64_t* values_out, uint32_t* status_out, double* latency_us_out, double* makespan_us);

/* Server over SingleValueHashTable<uint32_t, uint32_t> */
WC_API void* wc_server_u32u32_create(const uint64_t* table_capacities, uint64_t num_tables, uint64_t max_requests, uint64_t item_keys, uint64_t max_step_keys, float max_load, int use_graph, int blocks_per_cu, uint64_t direct_keys);
WC_API void wc_server_u32u32_destroy(void* server);
WC_API int64_t wc_server_u32u32_submit(void* server, uint32_t op, uint32_t table, const uint32_t* keys_in, const uint32_t* values_in, uint32_t* values_out, uint32_t* status_out, uint64_t num_in);
WC_API int wc_server_u32u32_wait(void* server, int64_t id);
WC_API int wc_server_u32u32_drain(void* server);
WC_API double wc_server_u32u32_latency_us(void* server, int64_t id);
WC_API uint64_t wc_server_u32u32_step_of(void* server, int64_t id);
WC_API int wc_server_u32u32_stats(void* server, uint64_t* out3);
WC_API uint64_t wc_server_u32u32_direct_steps(void* server);
WC_API int wc_server_u32u32_counters(void* server, uint64_t* out4);
WC_API uint64_t wc_server_u32u32_table_reserved(void* server, uint32_t table);
WC_API uint64_t wc_server_u32u32_table_budget(void* server, uint32_t table);
WC_API int wc_server_u32u32_fail_step(void* server, uint64_t step);
WC_API uint64_t wc_server_u32u32_table_size(void* server, uint32_t table);
WC_API uint64_t wc_server_u32u32_table_capacity(void* server, uint32_t table);
WC_API uint32_t wc_server_u32u32_table_status(void* server, uint32_t table);
WC_API int wc_server_u32u32_drive(void* server, int mode, const uint32_t* ops, const uint32_t* tables, const uint64_t* sizes, const double* arrival_us, uint64_t num_requests, const uint32_t* keys_in, const uint32_t* values_in, uint32_t* values_out, uint32_t* status_out, double* latency_us_out, double* makespan_us);

/* ------------------------------------------------------------------------
 * Synthetic inputs (the role of include/warpcore/random_distributions.cuh:
 * uniform 50-135, unique 144-246, zipf 259-326). Counter-based: element i
 * depends only on (seed or offset, i).
 *   _unique_*   out[i] = fmix(offset + i), never 0 or ~0, all distinct
 *   _uniform_*  out[i] = splitmix64(seed, i)
 *   _zipf_*     Zipf(s) ranks over num_distinct keys (rejection-inversion,
 *               any s > 0), rank k -> unique key (key_offset, k - 1)
 *   _interleave_* out[2j], out[2j+1] = a[j], b[j] in hashed order
 * ---------------------------------------------------------------------- */
WC_API int wc_gen_unique_u32(uint32_t* out, uint64_t n, uint64_t offset, void* stream);
WC_API int wc_gen_unique_u64(uint64_t* out, uint64_t n, uint64_t offset, void* stream);
WC_API int wc_gen_uniform_u32(uint32_t* out, uint64_t n, uint64_t seed, void* stream);
WC_API int wc_gen_uniform_u64(uint64_t* out, uint64_t n, uint64_t seed, void* stream);
WC_API int wc_gen_zipf_u32(uint32_t* out, uint64_t n, uint64_t num_distinct, double s, uint64_t seed, uint64_t key_offset, void* stream);
WC_API int wc_gen_zipf_u64(uint64_t* out, uint64_t n, uint64_t num_distinct, double s, uint64_t seed, uint64_t key_offset, void* stream);
WC_API int wc_gen_interleave_u32(uint32_t* out, const uint32_t* a, const uint32_t* b, uint64_t half, uint64_t seed, void* stream);
WC_API int wc_gen_interleave_u64(uint64_t* out, const uint64_t* a, const uint64_t* b, uint64_t half, uint64_t seed, void* stream);

/* Key files: raw little-endian arrays of 4- or 8-byte keys, the input format
 * of the reference benchmarks' load_keys / helpers::load_binary
 * (benchmarks/include/benchmark_common.cuh:62-75). At most max_keys keys
 * are read from the start of the file (fewer if it is shorter).
 *   _load_binary_*  into a HOST buffer of max_keys keys; returns the count
 *                   read, or -1 when the file cannot be read
 *   _load_keys_*    into a DEVICE buffer of max_keys keys; *num_out = count
 *                   read; returns 0, -2 (unreadable file / bad argument) or
 *                   a hipError_t. Unlike the reference, only the keys read
 *                   are copied. */
WC_API int64_t wc_load_binary_u32(const char* path, uint64_t max_keys, uint32_t* keys_out);
WC_API int64_t wc_load_binary_u64(const char* path, uint64_t max_keys, uint64_t* keys_out);
WC_API int wc_load_keys_u32(const char* path, uint64_t max_keys, uint32_t* keys_d, uint64_t* num_out, void* stream);
WC_API int wc_load_keys_u64(const char* path, uint64_t max_keys, uint64_t* keys_d, uint64_t* num_out, void* stream);

/* Capacity rule detail::get_valid_capacity (include/warpcore/base.cuh:27-34)
 * and the unique-key map, evaluated on the host. */
WC_API uint64_t wc_get_valid_capacity(uint64_t min_capacity, uint64_t cg_size);
WC_API uint64_t wc_unique_key_u64(uint64_t offset, uint64_t i);
WC_API uint32_t wc_unique_key_u32(uint64_t offset, uint64_t i);

/* Kernel-shape policy (no reference counterpart): bulk lookups on tables
 * whose slots take at most `bytes` (default 128 MiB) are probed one lane per
 * key, larger tables by cooperative groups of cg_size lanes per key; bulk
 * inserts likewise with their own limit (default 0: always groups). */
WC_API void wc_set_lane_table_bytes(uint64_t bytes);
WC_API uint64_t wc_get_lane_table_bytes(void);
WC_API void wc_set_lane_insert_table_bytes(uint64_t bytes);
WC_API uint64_t wc_get_lane_insert_table_bytes(void);

/* Bulk kernels of tables probed by cooperative groups (no reference
 * counterpart), a bit mask: bit 7 = lane-owner inserts, bit 6 = lane-owner
 * lookups and erases (single-value tables with windows of 2..16 slots;
 * these win over the rest), bit 8 = lane-owner kernels read packed 8-byte
 * slots two per lane, bit 3 = refill inserts, bit 2 = refill lookups,
 * bit 4 = pipelined refill lookups, bit 0 = lockstep inserts, bit 1 =
 * lockstep lookups and erases (a refill bit wins over the lockstep one; no
 * bit = the streaming per-key state machine). Default 476 (bits 2, 3, 4, 6,
 * 7, 8). MultiValueHashTable bulk calls use the refill kernels under bits 3
 * (inserts) and 2 (count and gather passes), else one key per group. All
 * choices give identical results. */
WC_API void wc_set_group_kernel(int kernel);
WC_API int wc_get_group_kernel(void);

/* CountingHashTable bulk inserts of at least 4096 keys (no reference
 * counterpart): 1 = equal keys of each tile of 1024 occurrences are combined
 * in LDS and inserted once with their count (default), 0 = every occurrence
 * inserted on its own. Counts and statuses are the same either way. */
WC_API void wc_set_counting_combine(int mode);
WC_API int wc_get_counting_combine(void);

#endif /* WARPCORE_C_H */
