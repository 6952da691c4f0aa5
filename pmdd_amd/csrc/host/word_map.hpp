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
        std::fill(slots_.begin(), slots_.end(), -1);
    }

    // Returns the id of `k`, inserting it when absent; *inserted tells which.
    int64_t insert(const uint64_t *k, bool *inserted) {
        if ((size() + 1) * 2 > static_cast<int64_t>(slots_.size())) {
            rehash(slots_.size() * 2);
        }
        const uint64_t h = hash(k);
        size_t mask = slots_.size() - 1;
        size_t pos = static_cast<size_t>(h) & mask;
        while (true) {
            const int64_t id = slots_[pos];
            if (id < 0) {
                const int64_t nid = size();
                keys_.insert(keys_.end(), k, k + words_);
                slots_[pos] = nid;
                *inserted = true;
                return nid;
            }
            if (std::memcmp(key(id), k, sizeof(uint64_t) * words_) == 0) {
                *inserted = false;
                return id;
            }
            pos = (pos + 1) & mask;
        }
    }

    int64_t find(const uint64_t *k) const {
        const uint64_t h = hash(k);
        size_t mask = slots_.size() - 1;
        size_t pos = static_cast<size_t>(h) & mask;
        while (true) {
            const int64_t id = slots_[pos];
            if (id < 0) {
                return -1;
            }
            if (std::memcmp(key(id), k, sizeof(uint64_t) * words_) == 0) {
                return id;
            }
            pos = (pos + 1) & mask;
        }
    }

  private:
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

    void rehash(size_t n) {
        slots_.assign(n, -1);
        const size_t mask = n - 1;
        const int64_t count = size();
        for (int64_t id = 0; id < count; ++id) {
            size_t pos = static_cast<size_t>(hash(key(id))) & mask;
            while (slots_[pos] >= 0) {
                pos = (pos + 1) & mask;
            }
            slots_[pos] = id;
        }
    }

    int words_;
    std::vector<uint64_t> keys_;
    std::vector<int64_t> slots_;
};

} // namespace pmdd
