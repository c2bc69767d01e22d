// User kernels that call the tables' device API (the tile-based members),
// as a reference user would (examples/src/advanced_usage_from_device.cu:14-66,
// examples/src/multi_value_hash_table.cu). Test infrastructure: built into
// tests/device_api/libwc_device_api.so and driven from the GPU tests.
#include <warpcore/warpcore.hpp>

using namespace warpcore;

namespace {

// filtered histogram: keys seen before (Bloom insert_and_query) are counted
template<class Key, class Filter, class Table>
__global__ void filter_and_count(const Key* keys, index_t n, Filter filter, Table table) {
    const index_t tid = index_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const tile<Table::cg_size()> ht_group;
    const tile<Filter::cg_size()> bf_group;
    Key key{};
    bool seen = false;
    if (tid < n) {
        key = keys[tid];
        seen = filter.insert_and_query(key, bf_group);
    }
    auto mask = ht_group.ballot(seen);
    while (mask) {
        const unsigned leader = unsigned(ffs64(mask) - 1);
        const Key k = ht_group.shfl(key, leader);
        table.insert(k, ht_group);
        mask &= mask - 1;
    }
}

template<class Table>
__global__ void svht_insert_retrieve(const typename Table::key_type* keys, const typename Table::value_type* values,
                                     index_t n, Table table, int phase, typename Table::value_type* values_out,
                                     std::uint32_t* status_out) {
    constexpr unsigned CG = unsigned(Table::cg_size());
    const tile<CG> group;
    const index_t gid = (index_t(blockIdx.x) * blockDim.x + threadIdx.x) / CG;
    if (gid >= n) return;
    Status st;
    if (phase == 0) {
        st = table.insert(keys[gid], values[gid], group);
    } else if (phase == 1) {
        typename Table::value_type v{};
        st = table.retrieve(keys[gid], v, group);
        if (group.thread_rank() == 0 && !st.has_any()) values_out[gid] = v;
    } else {
        st = table.erase(keys[gid], group);
    }
    if (group.thread_rank() == 0) status_out[gid] = st.base();
}

template<class Table>
__global__ void mvht_insert(const typename Table::key_type* keys, const typename Table::value_type* values,
                            index_t n, Table table, std::uint32_t* status_out) {
    constexpr unsigned CG = unsigned(Table::cg_size());
    const tile<CG> group;
    const index_t gid = (index_t(blockIdx.x) * blockDim.x + threadIdx.x) / CG;
    if (gid >= n) return;
    const Status st = table.insert(keys[gid], values[gid], group);
    if (group.thread_rank() == 0) status_out[gid] = st.base();
}

// values of key i go to values_out[i * stride ...]; count to counts_out[i]
template<class Table>
__global__ void mvht_retrieve(const typename Table::key_type* keys, index_t n, Table table,
                              typename Table::value_type* values_out, index_t stride, index_t* counts_out,
                              std::uint32_t* status_out) {
    constexpr unsigned CG = unsigned(Table::cg_size());
    const tile<CG> group;
    const index_t gid = (index_t(blockIdx.x) * blockDim.x + threadIdx.x) / CG;
    if (gid >= n) return;
    index_t num = 0;
    const Status st = table.num_values(keys[gid], num, group);
    index_t got = 0;
    if (num <= stride) (void)table.retrieve(keys[gid], values_out + gid * stride, got, group);
    if (group.thread_rank() == 0) {
        counts_out[gid] = num;
        status_out[gid] = st.base();
    }
}

template<class Table>
__global__ void bucketlist_roundtrip(const typename Table::key_type* keys, const typename Table::value_type* values,
                                     index_t n, Table table, int phase, typename Table::value_type* values_out,
                                     index_t stride, index_t* counts_out, std::uint32_t* status_out) {
    constexpr unsigned CG = unsigned(Table::cg_size());
    const tile<CG> group;
    const index_t gid = (index_t(blockIdx.x) * blockDim.x + threadIdx.x) / CG;
    if (gid >= n) return;
    Status st;
    if (phase == 0) {
        st = table.insert(keys[gid], values[gid], group);
    } else {
        index_t num = 0;
        st = table.num_values(keys[gid], num, group);
        index_t got = 0;
        if (!st.has_any() && num <= stride) (void)table.retrieve(keys[gid], values_out + gid * stride, got, group);
        if (group.thread_rank() == 0) counts_out[gid] = num;
    }
    if (group.thread_rank() == 0) status_out[gid] = st.base();
}

unsigned blocks_for(index_t threads) { return unsigned((threads + 255) / 256); }

}  // namespace

extern "C" {

// Counts of the keys that occur more than once (count = occurrences - 1
// unless the filter reported a false positive). Returns the number of
// distinct keys written to keys_out/counts_out (device buffers of n).
int wcx_filter_and_count_u32(const std::uint32_t* keys, std::uint64_t n, std::uint64_t filter_bits,
                             std::uint32_t* keys_out, std::uint64_t* counts_out, std::uint64_t* num_out) {
    BloomFilter<std::uint32_t> filter(filter_bits, 6);
    CountingHashTable<std::uint32_t> table(n + 16);
    filter_and_count<<<blocks_for(n ? n : 1), 256>>>(keys, n, filter, table);
    index_t num = 0;
    table.retrieve_all(keys_out, counts_out, num);
    *num_out = num;
    return int(hipDeviceSynchronize());
}

// SingleValueHashTable<u64,u64> (cg 8): phase 0 insert, 1 retrieve, 2 erase,
// all from user kernels on one table that lives across the calls.
void* wcx_svht_create(std::uint64_t capacity) { return new SingleValueHashTable<std::uint64_t, std::uint64_t>(capacity); }
void wcx_svht_destroy(void* t) { delete static_cast<SingleValueHashTable<std::uint64_t, std::uint64_t>*>(t); }
int wcx_svht_phase(void* t, int phase, const std::uint64_t* keys, const std::uint64_t* values, std::uint64_t n,
                   std::uint64_t* values_out, std::uint32_t* status_out) {
    auto& table = *static_cast<SingleValueHashTable<std::uint64_t, std::uint64_t>*>(t);
    svht_insert_retrieve<<<blocks_for(n * table.cg_size() + 1), 256>>>(keys, values, n, table, phase, values_out,
                                                                         status_out);
    return int(hipDeviceSynchronize());
}

// MultiValueHashTable<u32,u32> (cg 8) through the device API.
int wcx_mvht_roundtrip_u32(const std::uint32_t* keys, const std::uint32_t* values, std::uint64_t n,
                           const std::uint32_t* query, std::uint64_t nq, std::uint64_t capacity,
                           std::uint32_t* insert_status, std::uint32_t* values_out, std::uint64_t stride,
                           std::uint64_t* counts_out, std::uint32_t* query_status) {
    MultiValueHashTable<std::uint32_t, std::uint32_t> table(capacity);
    mvht_insert<<<blocks_for(n * table.cg_size() + 1), 256>>>(keys, values, n, table, insert_status);
    mvht_retrieve<<<blocks_for(nq * table.cg_size() + 1), 256>>>(query, nq, table, values_out, stride, counts_out,
                                                                   query_status);
    return int(hipDeviceSynchronize());
}

// BucketListHashTable<u32,u32> (cg 8) through the device API.
int wcx_bucketlist_roundtrip_u32(const std::uint32_t* keys, const std::uint32_t* values, std::uint64_t n,
                                 const std::uint32_t* query, std::uint64_t nq, std::uint64_t key_capacity,
                                 std::uint64_t value_capacity, std::uint32_t* insert_status,
                                 std::uint32_t* values_out, std::uint64_t stride, std::uint64_t* counts_out,
                                 std::uint32_t* query_status) {
    BucketListHashTable<std::uint32_t, std::uint32_t> table(key_capacity, value_capacity);
    bucketlist_roundtrip<<<blocks_for(n * table.cg_size() + 1), 256>>>(keys, values, n, table, 0, nullptr, 0,
                                                                         nullptr, insert_status);
    bucketlist_roundtrip<<<blocks_for(nq * table.cg_size() + 1), 256>>>(query, nullptr, nq, table, 1, values_out,
                                                                          stride, counts_out, query_status);
    return int(hipDeviceSynchronize());
}

}  // extern "C"
