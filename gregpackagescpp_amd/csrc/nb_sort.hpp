// nb_sort.hpp -- device radix sort (nb_sort.hip).
#pragma once
#include <utility>

#include "nb_internal.hpp"

namespace nbe {

struct RadixSorter {
    static constexpr int kDigitBits = 9;  // bits per pass
    DBuf<uint64_t> k2;
    DBuf<uint32_t> v2;
    DBuf<uint32_t> hist;  // [digit][tile] counts, then in-digit tile offsets
    DBuf<uint32_t> dtot;  // per-digit totals
    void ensure(int64_t n);
    void sort(hipStream_t s, uint64_t *keys, uint32_t *vals, int64_t n, int bits);
    // by key bits [lo_bit, bits) only (stable, so equal such bits keep the input order)
    void sort_swap(hipStream_t s, DBuf<uint64_t> &keys, DBuf<uint32_t> &vals, int64_t n, int bits, int lo_bit = 0);
    std::pair<uint64_t *, uint32_t *> passes(hipStream_t s, uint64_t *keys, uint32_t *vals, int64_t n, int bits,
                                             int lo_bit = 0);
};

}  // namespace nbe
