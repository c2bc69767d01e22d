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

// Both boolean status handlers, for one compile-time Ignore word.
template<Status::base_type Ignore>
__global__ void status_handlers_kernel(const std::uint32_t* statuses, index_t n, bool* rb_out, bool* any_out) {
    const index_t i = index_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Status s = Status::from_base(statuses[i]);
    status_handlers::ReturnBoolean<Ignore>::handle(s, rb_out, i);
    status_handlers::ReturnAnyFlag<Ignore>::handle(s, any_out, i);
}

// Windows a probing scheme visits for each key: one group of CG lanes per
// key walks begin()/next() and lane 0 writes slot / CG per probe.
template<class Scheme, class K>
__global__ void probe_windows_kernel(const K* keys, index_t n, K seed, probing_schemes::Geometry geo, index_t pl,
                                     index_t max_windows, std::uint64_t* windows_out, std::uint32_t* counts_out) {
    constexpr unsigned CG = unsigned(Scheme::cg_size());
    const tile<CG> group;
    const index_t gid = (index_t(blockIdx.x) * blockDim.x + threadIdx.x) / CG;
    if (gid >= n) return;
    Scheme iter(geo, pl, group);
    index_t w = 0;
    for (index_t i = iter.begin(keys[gid], seed); i != iter.end() && w < max_windows; i = iter.next(), ++w) {
        if (group.thread_rank() == 0) windows_out[gid * max_windows + w] = i / CG;
        if (i % CG != group.thread_rank()) windows_out[gid * max_windows + w] = ~std::uint64_t(0);   // not aligned
    }
    if (group.thread_rank() == 0) counts_out[gid] = std::uint32_t(w);
}

// The same walk with the window cursor of the bulk kernels.
template<class Scheme, class K>
__global__ void cursor_windows_kernel(const K* keys, index_t n, K seed, probing_schemes::Geometry geo, index_t pl,
                                      index_t max_windows, std::uint64_t* windows_out, std::uint32_t* counts_out) {
    const index_t i = index_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    detail::WindowCursor<Scheme> c;
    c.begin(keys[i], seed, geo);
    c.left = probing_schemes::detail::window_budget(pl, Scheme::cg_size(), geo.windows);
    index_t w = 0;
    do {
        windows_out[i * max_windows + w] = c.w;
        ++w;
    } while (w < max_windows && c.advance(geo));
    counts_out[i] = std::uint32_t(w);
}

template<class Scheme, class K>
int run_probe_windows(bool cursor, const void* keys, index_t n, std::uint64_t seed, index_t capacity, index_t pl,
                      index_t max_windows, std::uint64_t* windows_out, std::uint32_t* counts_out) {
    const auto geo = probing_schemes::Geometry::make(capacity, Scheme::cg_size());
    if (cursor)
        cursor_windows_kernel<Scheme, K><<<unsigned((n + 255) / 256), 256>>>(
            static_cast<const K*>(keys), n, K(seed), geo, pl, max_windows, windows_out, counts_out);
    else
        probe_windows_kernel<Scheme, K><<<unsigned((n * Scheme::cg_size() + 255) / 256), 256>>>(
            static_cast<const K*>(keys), n, K(seed), geo, pl, max_windows, windows_out, counts_out);
    return int(hipDeviceSynchronize());
}

template<class K, class H>
int probe_windows_for(int scheme, const void* keys, index_t n, std::uint64_t seed, index_t capacity, index_t pl,
                      index_t max_windows, std::uint64_t* windows_out, std::uint32_t* counts_out) {
    using namespace probing_schemes;
    const bool cursor = (scheme & 16) != 0;
    switch (scheme & 15) {
    case 0: return run_probe_windows<DoubleHashing<H, H, 8>, K>(cursor, keys, n, seed, capacity, pl, max_windows, windows_out, counts_out);
    case 1: return run_probe_windows<LinearProbing<H, 8>, K>(cursor, keys, n, seed, capacity, pl, max_windows, windows_out, counts_out);
    case 2: return run_probe_windows<QuadraticProbing<H, 8>, K>(cursor, keys, n, seed, capacity, pl, max_windows, windows_out, counts_out);
    default: return -1;
    }
}

// tests/src/key_value_store_test.cu:26-160 for one store type: errors per
// section in errors[0..4] (constructor, set/get pairs, init keys, init
// pairs, key CAS leaves the value). errors must be zeroed device words.
template<class Store, class Key, class Value>
__global__ void store_set(Store st, index_t n) {
    const index_t i = index_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        auto&& p = st[i];
        p.key = Key(i);
        p.value = Value(i);
    }
}
template<class Store, class Key, class Value>
__global__ void store_get(Store st, index_t n, unsigned long long* errors) {
    const index_t i = index_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n && (st[i].key != Key(i) || st[i].value != Value(i))) atomicAdd(errors, 1ull);
}
template<class Store, class Key, class Value>
__global__ void store_check_pairs(Store st, index_t n, Key key, Value value, bool check_value,
                                  unsigned long long* errors) {
    const index_t i = index_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n && (st[i].key != key || (check_value && st[i].value != value))) atomicAdd(errors, 1ull);
}
template<class Store, class Key, class Value>
__global__ void store_cas(Store st, Key init, Key key, Value value, unsigned long long* errors) {
    atomicCAS(&st[0].key, init, key);
    if (!(st[0].key == key && st[0].value == value)) atomicAdd(errors, 1ull);
}

// Sets the count of `key` (a slot of the counting table) to `count`.
template<class View>
__global__ void counting_preset(View v, typename View::key_type key, typename View::value_type count) {
    const index_t i = index_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < v.capacity() && v.base.slots[i].key == key) v.base.slots[i].value = count;
}

template<class Store>
int store_check(std::uint64_t capacity, unsigned long long* errors) {
    using Key = typename Store::key_type;
    using Value = typename Store::value_type;
    {
        Store st(capacity);
        if (st.status() != Status::none() || st.capacity() != capacity) {
            const unsigned long long one = 1;
            (void)hipMemcpy(errors + 0, &one, sizeof(one), hipMemcpyHostToDevice);
        }
        store_set<Store, Key, Value><<<blocks_for(capacity), 256>>>(st, capacity);
        store_get<Store, Key, Value><<<blocks_for(capacity), 256>>>(st, capacity, errors + 1);
        st.init_keys(Key(42));
        store_check_pairs<Store, Key, Value><<<blocks_for(capacity), 256>>>(st, capacity, Key(42), Value(0), false,
                                                                            errors + 2);
        st.init_pairs(Key(42), Value(1337));
        store_check_pairs<Store, Key, Value><<<blocks_for(capacity), 256>>>(st, capacity, Key(42), Value(1337), true,
                                                                            errors + 3);
        if (hipDeviceSynchronize() != hipSuccess) return -1;
    }
    Store one(1);
    one.init_pairs(Key(0), Value(1337));
    store_cas<Store, Key, Value><<<1, 1>>>(one, Key(0), Key(42), Value(1337), errors + 4);
    return int(hipDeviceSynchronize());
}



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

// Flags of ReturnBoolean<I> and ReturnAnyFlag<I> for the statuses, with
// I = none (which 0), duplicate_key (1), duplicate_key + key_not_found (2),
// invalid_value (3).
int wcx_status_handlers(const std::uint32_t* statuses, std::uint64_t n, int which, bool* rb_out, bool* any_out) {
    const unsigned blocks = blocks_for(n ? n : 1);
    switch (which) {
    case 0: status_handlers_kernel<0><<<blocks, 256>>>(statuses, n, rb_out, any_out); break;
    case 1: status_handlers_kernel<(1u << 4)><<<blocks, 256>>>(statuses, n, rb_out, any_out); break;
    case 2: status_handlers_kernel<(1u << 4) | (1u << 5)><<<blocks, 256>>>(statuses, n, rb_out, any_out); break;
    case 3: status_handlers_kernel<(1u << 12)><<<blocks, 256>>>(statuses, n, rb_out, any_out); break;
    default: return -1;
    }
    return int(hipDeviceSynchronize());
}

// Window sequences of the device probing schemes (cg 8): scheme 0 double,
// 1 linear, 2 quadratic, plus 16 to walk detail::WindowCursor (the bulk
// kernels' cursor) instead of the iterator class; hasher 0 MurmurHash, 1 MuellerHash, 2 NvidiaHash,
// 3 IdentityMap (Mueller and Nvidia for 32-bit keys only). windows_out is
// n x max_windows, counts_out[i] the number of windows key i visited.
int wcx_probe_windows(int scheme, int hasher, int key_bits, const void* keys, std::uint64_t n, std::uint64_t seed,
                      std::uint64_t capacity, std::uint64_t probing_length, std::uint64_t max_windows,
                      std::uint64_t* windows_out, std::uint32_t* counts_out) {
    using std::uint32_t;
    using std::uint64_t;
    if (key_bits == 32) {
        switch (hasher) {
        case 0: return probe_windows_for<uint32_t, hashers::MurmurHash<uint32_t>>(scheme, keys, n, seed, capacity, probing_length, max_windows, windows_out, counts_out);
        case 1: return probe_windows_for<uint32_t, hashers::MuellerHash>(scheme, keys, n, seed, capacity, probing_length, max_windows, windows_out, counts_out);
        case 2: return probe_windows_for<uint32_t, hashers::NvidiaHash>(scheme, keys, n, seed, capacity, probing_length, max_windows, windows_out, counts_out);
        case 3: return probe_windows_for<uint32_t, hashers::IdentityMap<uint32_t, uint32_t>>(scheme, keys, n, seed, capacity, probing_length, max_windows, windows_out, counts_out);
        default: return -1;
        }
    }
    switch (hasher) {
    case 0: return probe_windows_for<uint64_t, hashers::MurmurHash<uint64_t>>(scheme, keys, n, seed, capacity, probing_length, max_windows, windows_out, counts_out);
    case 3: return probe_windows_for<uint64_t, hashers::IdentityMap<uint64_t, uint64_t>>(scheme, keys, n, seed, capacity, probing_length, max_windows, windows_out, counts_out);
    default: return -1;
    }
}

// Store invariants: kind 0 AoSStore<u32,u32>, 1 AoSStore<u32,float>,
// 2 SoAStore<u32,u32>, 3 SoAStore<u32,float>.
int wcx_store_check(int kind, std::uint64_t capacity, unsigned long long* errors) {
    using namespace storage::key_value;
    switch (kind) {
    case 0: return store_check<AoSStore<std::uint32_t, std::uint32_t>>(capacity, errors);
    case 1: return store_check<AoSStore<std::uint32_t, float>>(capacity, errors);
    case 2: return store_check<SoAStore<std::uint32_t, std::uint32_t>>(capacity, errors);
    case 3: return store_check<SoAStore<std::uint32_t, float>>(capacity, errors);
    default: return -1;
    }
}

// CountingHashTable<u32, u32 counts> with max_count: key keys[0] is inserted
// once and its count set to `preset`, then the whole batch keys[0..n) is
// inserted by the host bulk call (combined counts for n >= 4096). Writes
// the per-occurrence statuses, the counts of keys[0..n) and the table
// status word.
int wcx_counting_u32_saturation(const std::uint32_t* keys, std::uint64_t n, std::uint32_t max_count,
                                std::uint32_t preset, std::uint32_t* status_out, std::uint32_t* counts_out,
                                std::uint32_t* table_status) {
    using table_t = CountingHashTable<std::uint32_t, std::uint32_t>;
    table_t table(2 * n + 1024, defaults::seed<std::uint32_t>(), max_count);
    table.insert(keys, 1);
    if (hipDeviceSynchronize() != hipSuccess) return -1;
    std::uint32_t first = 0;
    if (hipMemcpy(&first, keys, sizeof(first), hipMemcpyDeviceToHost) != hipSuccess) return -1;
    counting_preset<<<blocks_for(table.capacity()), 256>>>(table.view(), first, preset);
    table.template insert<status_handlers::ReturnStatus>(keys, n, 0, defaults::probing_length(),
                                                          reinterpret_cast<Status*>(status_out));
    table.retrieve(keys, n, counts_out);
    *table_status = table.peek_status().base();
    return int(hipDeviceSynchronize());
}

}  // extern "C"
