// Slot storage for the key/value tables and the temporary-word ring.
//
// AoSStore keeps pair_t{key, value} slots like the reference
// (include/warpcore/storage.cuh:438-645). Slots of 8 or 16 bytes are aligned
// to their size so one lane reads a whole slot with one dwordx2/dwordx4 load
// and a window of CG slots is one contiguous, line-aligned span; an 8-byte
// u32/u32 slot is also the unit of the packed 64-bit compare-and-swap.
// SoAStore keeps keys and values in two arrays (storage.cuh:218-432): a
// window of CG keys is then CG * sizeof(Key) contiguous bytes and a value is
// read only for the slot that holds the key.
// CyclicStore is the ring of device temp words that host calls borrow for
// counters (storage.cuh:15-148).
#pragma once

#include "base.hpp"
#include "status.hpp"
#include "tags.hpp"

#include <atomic>

namespace warpcore {
namespace storage {

namespace detail {
template<class K, class V> struct plain_pair { K key; V value; };
template<class K, class V>
constexpr std::size_t pair_alignment() noexcept {
    constexpr std::size_t s = sizeof(plain_pair<K, V>);
    return (s == 8 || s == 16) ? s : alignof(plain_pair<K, V>);
}
}  // namespace detail

// Temporary device words handed out round-robin (storage.cuh:15-148).
template<class T>
class CyclicStore {
public:
    using base_type = T;
    using index_type = index_t;
    using status_type = Status;

    WC_H explicit CyclicStore(index_type capacity) noexcept
        : store_(nullptr), capacity_(capacity), current_(nullptr), status_(Status::not_initialized()), is_copy_(false) {
        if (capacity_ == 0) { status_ += Status::invalid_configuration(); return; }
        if (warpcore::detail::available_gpu_memory() < sizeof(T) * capacity_ ||
            hipMalloc(&store_, sizeof(T) * capacity_) != hipSuccess) {
            store_ = nullptr;
            status_ += Status::out_of_memory();
            return;
        }
        current_ = new std::atomic<index_type>(0);
        status_ = Status::none();
    }
    WC_HD CyclicStore(const CyclicStore& o) noexcept
        : store_(o.store_), capacity_(o.capacity_), current_(o.current_), status_(o.status_), is_copy_(true) {}
    WC_H CyclicStore(CyclicStore&& o) noexcept
        : store_(o.store_), capacity_(o.capacity_), current_(o.current_), status_(o.status_), is_copy_(o.is_copy_) {
        o.is_copy_ = true;
    }
    WC_HD ~CyclicStore() noexcept {
#if !defined(__HIP_DEVICE_COMPILE__)
        if (!is_copy_) {
            if (store_ != nullptr) (void)hipFree(store_);
            delete current_;
        }
#endif
    }

    // Next word of the ring; concurrent host callers get distinct words.
    WC_H T* get() const noexcept {
        if (current_ == nullptr) return nullptr;
        index_type old = current_->load();
        while (!current_->compare_exchange_weak(old, old + 1 == capacity_ ? 0 : old + 1)) {}
        return store_ + old;
    }
    WC_HD status_type status() const noexcept { return status_; }
    WC_HD index_type capacity() const noexcept { return capacity_; }
    WC_H index_type bytes_total() const noexcept { return capacity_ * sizeof(T) + sizeof(index_type); }

private:
    T* store_;
    index_type capacity_;
    std::atomic<index_type>* current_;
    status_type status_;
    bool is_copy_;
};

namespace key_value {

template<class Key, class Value>
struct alignas(detail::pair_alignment<Key, Value>()) pair_t {
    Key key;
    Value value;
};

// References to the key and value of one SoA slot (storage.cuh:189-210).
template<class Key, class Value>
struct pair_ref_t {
    Key& key;
    Value& value;
};

// Struct-of-arrays slots (storage.cuh:218-432).
template<class Key, class Value>
class SoAStore {
public:
    using key_type = Key;
    using value_type = Value;
    using slot_type = pair_t<Key, Value>;   // what a slot reads as (by value)
    using status_type = Status;
    using index_type = index_t;
    using tag = tags::key_value_storage;

    WC_H explicit SoAStore(index_type capacity) noexcept
        : status_(Status::not_initialized()), capacity_(capacity), keys_(nullptr), values_(nullptr), is_copy_(false) {
        if (capacity_ == 0) { status_ += Status::invalid_configuration(); return; }
        const std::size_t bytes = (sizeof(Key) + sizeof(Value)) * capacity_;
        if (warpcore::detail::available_gpu_memory() < bytes || hipMalloc(&keys_, sizeof(Key) * capacity_) != hipSuccess ||
            hipMalloc(&values_, sizeof(Value) * capacity_) != hipSuccess) {
            release();
            status_ += Status::out_of_memory();
            return;
        }
        status_ = Status::none();
    }
    WC_HD SoAStore(const SoAStore& o) noexcept
        : status_(o.status_), capacity_(o.capacity_), keys_(o.keys_), values_(o.values_), is_copy_(true) {}
    WC_H SoAStore(SoAStore&& o) noexcept
        : status_(o.status_), capacity_(o.capacity_), keys_(o.keys_), values_(o.values_), is_copy_(o.is_copy_) {
        o.is_copy_ = true;
    }
    WC_HD ~SoAStore() noexcept {
#if !defined(__HIP_DEVICE_COMPILE__)
        if (!is_copy_) release();
#endif
    }

    WC_H void init_keys(key_type key, hipStream_t stream = 0) noexcept;
    WC_H void init_values(value_type value, hipStream_t stream = 0) noexcept;
    WC_H void init_pairs(key_type key, value_type value, hipStream_t stream = 0) noexcept {
        init_keys(key, stream);
        init_values(value, stream);
    }

    WC_D pair_ref_t<key_type, value_type> operator[](index_type i) noexcept { return {keys_[i], values_[i]}; }
    WC_D pair_ref_t<const key_type, const value_type> operator[](index_type i) const noexcept {
        return {keys_[i], values_[i]};
    }

    WC_HD key_type* keys() const noexcept { return keys_; }
    WC_HD value_type* values() const noexcept { return values_; }
    WC_HD status_type status() const noexcept { return status_; }
    WC_HD index_type capacity() const noexcept { return capacity_; }
    WC_H index_type bytes_total() const noexcept { return capacity_ * (sizeof(key_type) + sizeof(value_type)); }

private:
    WC_H void release() noexcept {
        if (keys_ != nullptr) (void)hipFree(keys_);
        if (values_ != nullptr) (void)hipFree(values_);
        keys_ = nullptr;
        values_ = nullptr;
    }

    status_type status_;
    index_type capacity_;
    key_type* keys_;
    value_type* values_;
    bool is_copy_;
};

// Array-of-structs slots (storage.cuh:438-645).
template<class Key, class Value>
class AoSStore {
public:
    using key_type = Key;
    using value_type = Value;
    using slot_type = pair_t<Key, Value>;
    using status_type = Status;
    using index_type = index_t;
    using tag = tags::key_value_storage;

    WC_H explicit AoSStore(index_type capacity) noexcept
        : status_(Status::not_initialized()), capacity_(capacity), pairs_(nullptr), is_copy_(false) {
        if (capacity_ == 0) { status_ += Status::invalid_configuration(); return; }
        const std::size_t bytes = sizeof(slot_type) * capacity_;
        if (warpcore::detail::available_gpu_memory() < bytes || hipMalloc(&pairs_, bytes) != hipSuccess) {
            pairs_ = nullptr;
            status_ += Status::out_of_memory();
            return;
        }
        status_ = Status::none();
    }
    WC_HD AoSStore(const AoSStore& o) noexcept
        : status_(o.status_), capacity_(o.capacity_), pairs_(o.pairs_), is_copy_(true) {}
    WC_H AoSStore(AoSStore&& o) noexcept
        : status_(o.status_), capacity_(o.capacity_), pairs_(o.pairs_), is_copy_(o.is_copy_) {
        o.is_copy_ = true;
    }
    WC_HD ~AoSStore() noexcept {
#if !defined(__HIP_DEVICE_COMPILE__)
        if (!is_copy_ && pairs_ != nullptr) (void)hipFree(pairs_);
#endif
    }

    // Write EMPTY keys (and `value`) into every slot with full-slot stores.
    WC_H void init_pairs(key_type key, value_type value, hipStream_t stream = 0) noexcept;
    WC_H void init_keys(key_type key, hipStream_t stream = 0) noexcept { init_pairs(key, value_type{}, stream); }
    WC_H void init_values(value_type value, hipStream_t stream = 0) noexcept;

    WC_D slot_type& operator[](index_type i) noexcept { return pairs_[i]; }
    WC_D const slot_type& operator[](index_type i) const noexcept { return pairs_[i]; }

    WC_HD slot_type* data() const noexcept { return pairs_; }
    WC_HD status_type status() const noexcept { return status_; }
    WC_HD index_type capacity() const noexcept { return capacity_; }
    WC_H index_type bytes_total() const noexcept { return capacity_ * sizeof(slot_type); }

private:
    status_type status_;
    index_type capacity_;
    slot_type* pairs_;
    bool is_copy_;
};

}  // namespace key_value

namespace kernels {

template<class Slot, class Key, class Value>
__global__ __launch_bounds__(256) void fill_pairs(Slot* __restrict__ slots, index_t n, Key key, Value value) {
    const index_t stride = index_t(gridDim.x) * blockDim.x;
    for (index_t i = index_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        Slot s;
        s.key = key;
        s.value = value;
        slots[i] = s;
    }
}

template<class Slot, class Value>
__global__ __launch_bounds__(256) void fill_values(Slot* __restrict__ slots, index_t n, Value value) {
    const index_t stride = index_t(gridDim.x) * blockDim.x;
    for (index_t i = index_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) slots[i].value = value;
}

template<class T>
__global__ __launch_bounds__(256) void fill(T* __restrict__ arr, index_t n, T value) {
    const index_t stride = index_t(gridDim.x) * blockDim.x;
    for (index_t i = index_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) arr[i] = value;
}

}  // namespace kernels

namespace key_value {

template<class Key, class Value>
WC_H void AoSStore<Key, Value>::init_pairs(key_type key, value_type value, hipStream_t stream) noexcept {
    if (status_.has_any() || capacity_ == 0) return;
    kernels::fill_pairs<<<warpcore::detail::grid_for(capacity_, 1, 8), block_size(), 0, stream>>>(pairs_, capacity_, key, value);
}

template<class Key, class Value>
WC_H void AoSStore<Key, Value>::init_values(value_type value, hipStream_t stream) noexcept {
    if (status_.has_any() || capacity_ == 0) return;
    kernels::fill_values<<<warpcore::detail::grid_for(capacity_, 1, 8), block_size(), 0, stream>>>(pairs_, capacity_, value);
}

template<class Key, class Value>
WC_H void SoAStore<Key, Value>::init_keys(key_type key, hipStream_t stream) noexcept {
    if (status_.has_any() || capacity_ == 0) return;
    kernels::fill<<<warpcore::detail::grid_for(capacity_, 1, 8), block_size(), 0, stream>>>(keys_, capacity_, key);
}

template<class Key, class Value>
WC_H void SoAStore<Key, Value>::init_values(value_type value, hipStream_t stream) noexcept {
    if (status_.has_any() || capacity_ == 0) return;
    kernels::fill<<<warpcore::detail::grid_for(capacity_, 1, 8), block_size(), 0, stream>>>(values_, capacity_, value);
}

}  // namespace key_value

}  // namespace storage
}  // namespace warpcore
