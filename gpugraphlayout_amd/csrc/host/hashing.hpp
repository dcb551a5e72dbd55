// Open-addressing hash tables for the loaders and generators (linear probing, power-of-two
// capacity at most half full, multiplicative hashing). Much faster and smaller than the
// node-based standard containers at 10^8 entries.
#ifndef GFA2_HASHING_HPP
#define GFA2_HASHING_HPP
#include <cstdint>
#include <vector>

namespace gfa2host
{
    inline uint64_t hash_slot(uint64_t key) { return (key * 0x9E3779B97F4A7C15ull) >> 17; }

    // Set of non-zero 64-bit keys.
    class KeySet
    {
    public:
        explicit KeySet(uint64_t expected)
        {
            uint64_t cap = 1024;
            while (cap < 2 * expected) cap <<= 1;
            slot_.assign(cap, 0);
            mask_ = cap - 1;
        }
        void prefetch(uint64_t key) const { __builtin_prefetch(&slot_[hash_slot(key) & mask_]); }
        // True if the key was not present (and is now).
        bool insert(uint64_t key)
        {
            for (uint64_t h = hash_slot(key);; ++h)
            {
                uint64_t& s = slot_[h & mask_];
                if (s == key) return false;
                if (s == 0)
                {
                    s = key;
                    return true;
                }
            }
        }

    private:
        std::vector<uint64_t> slot_;
        uint64_t mask_;
    };

    // Map from 32-bit ids to dense 32-bit ids, numbered by first insertion; grows.
    class IdMap
    {
    public:
        explicit IdMap(uint64_t expected = 1024) { rehash(expected); }
        void prefetch(uint32_t id) const { __builtin_prefetch(&slot_[hash_slot((uint64_t)id + 1) & mask_]); }
        // Returns the dense id of `id`, assigning the next one if it is new.
        uint32_t get_or_add(uint32_t id)
        {
            const uint64_t key = (uint64_t)id + 1;   // 0 marks an empty slot
            for (uint64_t h = hash_slot(key);; ++h)
            {
                Entry& e = slot_[h & mask_];
                if (e.key == key) return e.val;
                if (e.key == 0)
                {
                    e.key = key;
                    e.val = count_++;
                    if (2 * (uint64_t)count_ > slot_.size()) rehash(2 * (uint64_t)count_);
                    return count_ - 1;
                }
            }
        }
        uint32_t size() const { return count_; }

    private:
        struct Entry
        {
            uint64_t key;
            uint32_t val;
        };
        std::vector<Entry> slot_;
        uint64_t mask_ = 0;
        uint32_t count_ = 0;
        void rehash(uint64_t expected)
        {
            uint64_t cap = 1024;
            while (cap < 2 * expected) cap <<= 1;
            std::vector<Entry> old;
            old.swap(slot_);
            slot_.assign(cap, Entry{0, 0});
            mask_ = cap - 1;
            for (const Entry& e : old)
            {
                if (e.key == 0) continue;
                for (uint64_t h = hash_slot(e.key);; ++h)
                {
                    Entry& s = slot_[h & mask_];
                    if (s.key == 0)
                    {
                        s = e;
                        break;
                    }
                }
            }
        }
    };
}

#endif
