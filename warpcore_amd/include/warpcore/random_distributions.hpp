// Seeded device generators for synthetic inputs (the role of
// include/warpcore/random_distributions.cuh: uniform 50-135, unique
// 144-246, zipf 259-326). Counter-based, so element i depends only on
// (seed, i): the same call gives the same array on any grid.
//
//   unique:  key_i = fmix(offset + i) over the key width. fmix is a
//            bijection, so keys are distinct; the two inputs whose image is
//            EMPTY (0) or TOMBSTONE (~0) are replaced by fmix(offset - 1)
//            and fmix(offset - 2), which no other index produces.
//   uniform: splitmix64 of (seed, i), truncated to the value width.
//   zipf:    rank k in [1, N] with P(k) ~ k^-s by Hoermann and Derflinger's
//            rejection-inversion (ACM TOMACS 6(3), 1996), valid for every
//            s > 0 including s = 1 (the reference's sampler asserts s != 1);
//            rank k becomes the unique key fmix(key_offset + k - 1).
#pragma once

#include "base.hpp"
#include "hashers.hpp"

namespace warpcore {
namespace random {

WC_HD std::uint64_t splitmix64(std::uint64_t x) noexcept {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

// uniform double in [0, 1) from 53 random bits
WC_HD double unit_double(std::uint64_t bits) noexcept { return double(bits >> 11) * (1.0 / 9007199254740992.0); }

template<class K>
WC_HD K unique_key(std::uint64_t offset, std::uint64_t i) noexcept {
    const K x = K(offset + i);
    const K h = hashers::MurmurHash<K>::hash(x);
    if (h == K(0)) return hashers::MurmurHash<K>::hash(K(offset - 1));
    if (h == K(~K(0))) return hashers::MurmurHash<K>::hash(K(offset - 2));
    return h;
}

// Rejection-inversion sampler for Zipf(s) on {1..n}.
class ZipfSampler {
public:
    WC_HD ZipfSampler(std::uint64_t n, double s) noexcept : n_(double(n)), s_(s) {
        h_x1_ = H(1.5) - 1.0;
        h_n_ = H(n_ + 0.5);
        cut_ = 2.0 - H_inv(H(2.5) - h(2.0));
    }
    // draws uniforms from `next()` until a rank is accepted
    template<class Rng>
    WC_HD std::uint64_t operator()(Rng& next) const noexcept {
        while (true) {
            const double u = h_n_ + unit_double(next()) * (h_x1_ - h_n_);
            const double x = H_inv(u);
            double k = floor(x + 0.5);
            if (k < 1.0) k = 1.0;
            if (k > n_) k = n_;
            if (k - x <= cut_ || u >= H(k + 0.5) - h(k)) return std::uint64_t(k);
        }
    }

private:
    // H(x) = (x^(1-s) - 1) / (1-s), continuous at s = 1 (log x)
    WC_HD double H(double x) const noexcept {
        const double lx = log(x);
        const double t = (1.0 - s_) * lx;
        return (fabs(t) > 1e-8 ? expm1(t) / t : 1.0 + t * 0.5) * lx;
    }
    WC_HD double h(double x) const noexcept { return exp(-s_ * log(x)); }
    WC_HD double H_inv(double x) const noexcept {
        const double t = x * (1.0 - s_);
        const double f = fabs(t) > 1e-8 ? log1p(t) / t : 1.0 - t * 0.5;
        return exp(f * x);
    }
    double n_, s_, h_x1_, h_n_, cut_;
};

}  // namespace random

namespace kernels {

template<class K>
__global__ __launch_bounds__(256) void generate_unique(K* __restrict__ out, index_t n, std::uint64_t offset) {
    const index_t stride = index_t(gridDim.x) * blockDim.x;
    for (index_t i = index_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = random::unique_key<K>(offset, i);
}

template<class T>
__global__ __launch_bounds__(256) void generate_uniform(T* __restrict__ out, index_t n, std::uint64_t seed) {
    const index_t stride = index_t(gridDim.x) * blockDim.x;
    for (index_t i = index_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = T(random::splitmix64(seed ^ random::splitmix64(i)));
}

template<class K>
__global__ __launch_bounds__(256) void generate_zipf(K* __restrict__ out, index_t n, random::ZipfSampler zipf,
                                                     std::uint64_t seed, std::uint64_t key_offset) {
    const index_t stride = index_t(gridDim.x) * blockDim.x;
    for (index_t i = index_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        std::uint64_t ctr = random::splitmix64(seed ^ random::splitmix64(i));
        auto next = [&]() { ctr = random::splitmix64(ctr); return ctr; };
        const std::uint64_t k = zipf(next);
        out[i] = random::unique_key<K>(key_offset, k - 1);
    }
}

// Mixes two arrays of `half` keys into one stream of 2*half: positions 2j
// and 2j+1 hold a[j] and b[j], in an order picked by a hash of j.
template<class K>
__global__ __launch_bounds__(256) void interleave_halves(K* __restrict__ out, const K* __restrict__ a,
                                                         const K* __restrict__ b, index_t half, std::uint64_t seed) {
    const index_t stride = index_t(gridDim.x) * blockDim.x;
    for (index_t j = index_t(blockIdx.x) * blockDim.x + threadIdx.x; j < half; j += stride) {
        const bool swap = random::splitmix64(seed ^ j) & 1;
        out[2 * j] = swap ? b[j] : a[j];
        out[2 * j + 1] = swap ? a[j] : b[j];
    }
}

}  // namespace kernels
}  // namespace warpcore
