// Open-addressing hash map from fixed-width word keys to dense ids.
//
// Decision-diagram construction keys nodes by their state: used capacity per
// knapsack constraint, the eligible-vertex bitset of an independent-set node,
// or (visited set, last city) of a TSP node.  The reference uses
// boost::unordered_map for this (knapsack_bdd.hpp:18, indepset_bdd.hpp:31,
// tsp_mdd.cpp:40).  Ids are handed out in insertion order, so node order in a
// layer is the order in which states are first reached -- deterministic, and
// independent of any hash iteration order.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

namespace pmdd {

class WordKeyMap {
  public:
    explicit WordKeyMap(int words) : words_(words) { rehash(64); }

    int words() const { return words_; }
    int64_t size() const { return static_cast<int64_t>(keys_.size() / words_); }
    const uint64_t *key(int64_t id) const { return keys_.data() + id * words_; }

    void clear() {
        keys_.clear();
        std::fill(slots_.begin(), slots_.end(), kEmpty);
    }

    // Returns the id of `k`, inserting it when absent; *inserted tells which.
    int64_t insert(const uint64_t *k, bool *inserted) {
        reserve(size() + 1);
        return insert_hashed(k, hash(k), inserted);
    }

    // Inserts the n keys of `keys` (words() each, in order) and writes each
    // key's id to ids[i]: the same ids as n calls of insert(), with every
    // key's slot prefetched a few keys ahead (the slots are the cache misses).
    void insert_many(const uint64_t *keys, int64_t n, int64_t *ids) {
        constexpr int64_t kAhead = 16;
        reserve(size() + n);
        hashes_.resize(static_cast<size_t>(n));
        for (int64_t i = 0; i < n; ++i) {
            hashes_[i] = hash(keys + i * words_);
        }
        const size_t mask = slots_.size() - 1;
        for (int64_t i = 0; i < std::min(n, kAhead); ++i) {
            __builtin_prefetch(&slots_[static_cast<size_t>(hashes_[i]) & mask]);
        }
        bool inserted = false;
        for (int64_t i = 0; i < n; ++i) {
            if (i + kAhead < n) {
                __builtin_prefetch(&slots_[static_cast<size_t>(hashes_[i + kAhead]) & mask]);
            }
            ids[i] = insert_hashed(keys + i * words_, hashes_[i], &inserted);
        }
    }

    int64_t find(const uint64_t *k) const {
        const uint64_t h = hash(k);
        const uint64_t fp = h >> kIdBits;
        const size_t mask = slots_.size() - 1;
        size_t pos = static_cast<size_t>(h) & mask;
        while (true) {
            const uint64_t e = slots_[pos];
            if (e == kEmpty) {
                return -1;
            }
            const int64_t id = static_cast<int64_t>(e & kIdMask);
            if ((e >> kIdBits) == fp &&
                std::memcmp(key(id), k, sizeof(uint64_t) * words_) == 0) {
                return id;
            }
            pos = (pos + 1) & mask;
        }
    }

  private:
    // A slot holds the key's id in the low kIdBits bits and the top bits of
    // its hash above them, so most probes of other keys skip the key compare.
    static constexpr int kIdBits = 40;
    static constexpr uint64_t kIdMask = (uint64_t(1) << kIdBits) - 1;
    static constexpr uint64_t kEmpty = ~uint64_t(0);

    uint64_t hash(const uint64_t *k) const {
        // SplitMix64 finaliser folded over the words.
        uint64_t h = 0x9e3779b97f4a7c15ULL * static_cast<uint64_t>(words_);
        for (int i = 0; i < words_; ++i) {
            uint64_t z = k[i] + 0x9e3779b97f4a7c15ULL + (h << 6) + (h >> 2);
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
            h ^= z ^ (z >> 31);
        }
        return h;
    }

    // Room for `count` keys at a load of at most one half.
    void reserve(int64_t count) {
        if (count * 2 > static_cast<int64_t>(slots_.size())) {
            size_t n = slots_.size();
            while (static_cast<int64_t>(n) < count * 2) {
                n *= 2;
            }
            rehash(n);
        }
    }

    int64_t insert_hashed(const uint64_t *k, uint64_t h, bool *inserted) {
        const uint64_t fp = h >> kIdBits;
        const size_t mask = slots_.size() - 1;
        size_t pos = static_cast<size_t>(h) & mask;
        while (true) {
            const uint64_t e = slots_[pos];
            if (e == kEmpty) {
                const int64_t nid = size();
                keys_.insert(keys_.end(), k, k + words_);
                slots_[pos] = (fp << kIdBits) | static_cast<uint64_t>(nid);
                *inserted = true;
                return nid;
            }
            const int64_t id = static_cast<int64_t>(e & kIdMask);
            if ((e >> kIdBits) == fp &&
                std::memcmp(key(id), k, sizeof(uint64_t) * words_) == 0) {
                *inserted = false;
                return id;
            }
            pos = (pos + 1) & mask;
        }
    }

    void rehash(size_t n) {
        slots_.assign(n, kEmpty);
        const size_t mask = n - 1;
        const int64_t count = size();
        for (int64_t id = 0; id < count; ++id) {
            const uint64_t h = hash(key(id));
            size_t pos = static_cast<size_t>(h) & mask;
            while (slots_[pos] != kEmpty) {
                pos = (pos + 1) & mask;
            }
            slots_[pos] = ((h >> kIdBits) << kIdBits) | static_cast<uint64_t>(id);
        }
    }

    int words_;
    std::vector<uint64_t> keys_;
    std::vector<uint64_t> slots_;
    std::vector<uint64_t> hashes_;
};

} // namespace pmdd
