/*
 * warpcore_probes.h -- measurement probes of libwarpcore_probes.so (built
 * beside the table library for bench.py; not part of the table API).
 * Functions return 0, -2 for a bad argument, or the hipError_t of the
 * launch. `stream` is a hipStream_t passed as void*.
 */
#ifndef WARPCORE_PROBES_H
#define WARPCORE_PROBES_H

#include <stdint.h>

#ifdef __cplusplus
#define WC_API extern "C"
#else
#define WC_API
#endif

/* Device copy of `bytes` (a multiple of 16, 16-byte aligned pointers), 16
 * bytes per lane: a streaming kernel of known traffic, used to calibrate
 * memory-traffic counters. */
WC_API int wc_util_copy(void* dst, const void* src, uint64_t bytes, void* stream);
/* Reads `lines` random 128-byte lines of the first table_bytes bytes at
 * `table` (16-byte aligned), 8 lanes per line and 8 lines in flight per
 * group: the rate the memory system serves random whole lines at this
 * footprint, the practical ceiling of the probing kernels. `sink` is a
 * device word the kernel may write. */
WC_API int wc_util_gather_lines(const void* table, uint64_t table_bytes, uint64_t lines, uint64_t seed, uint32_t* sink, void* stream);
/* As wc_util_gather_lines, and then one lane per line replaces the 8 bytes
 * it read by their value plus one with a compare-and-swap that returns the
 * old word (the lane is chosen by the data read, so the CAS depends on the
 * line): the rate of "read a random line, claim a word of it", the memory
 * pattern of an insert without the probing. Writes into `table`. */
WC_API int wc_util_claim_lines(void* table, uint64_t table_bytes, uint64_t lines, uint64_t seed, uint32_t* sink, void* stream);

#endif
