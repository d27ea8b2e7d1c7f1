// bitgroups.h -- row groups for the bit-sliced multi-row kernels.
//
// Counting levels.  For counts ca, cb >= 0,
//     min(ca, cb) = #{ t >= 1 : t <= ca and t <= cb },
// so the intersection sum_x min(ca(x), cb(x)) of two Lingo multisets is the
// size of the intersection of their LEVEL sets {(x, t) : 1 <= t <= c(x)}.
//
// Level keys.  A Lingo id x whose four bytes are all below 0x80 (four ASCII
// characters, what the packer produces from SMILES) leaves bits 7, 15, 23
// and 31 clear; the level (x, t), 1 <= t <= 15, is encoded by placing the
// four bits of t there:
//     key(x, t) = x | (t&1) << 7 | (t&2) << 14 | (t&4) << 21 | (t&8) << 28.
// The map is injective on such (x, t) and never 0.  Levels with another x or
// with t > 15 cannot occur in any grouped row (see below), so the column
// stream omits them.
//
// Slots.  Every key has two candidate slots of a kBitSlots-slot table,
// h1(key) and h2(key), fixed functions of the key (a 32-bit mix), so a
// column's slot addresses are computed once when its level stream is built.
// A group is up to kBitRows reference rows whose level union has at most
// kBitMaxKeys keys; each key of the union sits in one of its two slots
// (cuckoo placement) holding (key, mask), bit r of mask saying that row r of
// the group has the level.  Empty slots hold (0, 0).  Probing a column level
// reads both candidate slots and takes the mask of the one whose key is
// equal; summing the masks as 32 vertical binary counters gives every row's
// intersection with the column at once.
//
// Rows qualify when their ids are strictly ascending and ASCII, every count
// is in [0, 15] and the row alone fits in a group; other rows use the merge
// kernel.
#pragma once

#include <cstdint>
#include <memory>
#include <vector>

namespace siml {

constexpr int kBitRows = 32;
constexpr int kBitRowStride = 32;  // row indices stored per group
constexpr int kBitSlotBits = 12;
constexpr int kBitSlots = 1 << kBitSlotBits;  // 4096 slots x 8 B = 32 KiB of LDS
constexpr int kBitMaxKeys = 2000;             // load factor <= 0.49
constexpr int kBitMaxLevel = 15;

inline bool ascii_id(int32_t x) { return ((uint32_t)x & 0x80808080u) == 0u; }

inline uint32_t level_key(int32_t x, uint32_t t) {
    return (uint32_t)x | ((t & 1u) << 7) | ((t & 2u) << 14) | ((t & 4u) << 21) | ((t & 8u) << 28);
}

// The two candidate slots of a key (distinct).
inline void key_slots(uint32_t key, uint32_t& h1, uint32_t& h2) {
    uint32_t h = key * 0x9E3779B1u;
    h ^= h >> 15;
    h *= 0x85EBCA77u;
    h ^= h >> 13;
    h *= 0xC2B2AE3Du;
    h ^= h >> 16;
    h1 = h >> (32 - kBitSlotBits);
    h2 = (h >> 4) & (kBitSlots - 1);
    if (h2 == h1) h2 ^= 1u;
}

// Second word of a column level: the byte offsets of both slots in the
// table (8-byte slots), h1 in the low half, h2 in the high half.
inline uint32_t slot_word(uint32_t key) {
    uint32_t h1, h2;
    key_slots(key, h1, h2);
    return (h1 * 8u) | ((h2 * 8u) << 16);
}

struct BitGroupHdr {
    int32_t nrows;      // 1..kBitRows
    int32_t nkeys;      // levels in the union
    int32_t nbits;      // bit planes the counts need: 8 or 16
    int32_t pad0;
    int64_t entry_off;  // first entry of the group in the entry array
};

struct BitEntry {
    uint32_t key;   // level key
    uint32_t mask;  // rows of the group holding the level
    int32_t slot;
    int32_t pad;
};

struct BitGroupInput {
    const int32_t* ids;
    const int32_t* counts;
    const int64_t* start;
    const int32_t* len;
    const int32_t* mag;  // groups are formed over rows sorted by magnitude when by_mag
    int64_t n;
};

// Allocator whose value-less construct leaves trivial elements
// uninitialised, so a resize that the caller overwrites at once does not
// first fill the memory with zeros.
template <class T>
struct NoInitAlloc : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = NoInitAlloc<U>;
    };
    NoInitAlloc() = default;
    template <class U>
    NoInitAlloc(const NoInitAlloc<U>&) {}
    template <class U>
    void construct(U* p) {
        ::new (static_cast<void*>(p)) U;
    }
    template <class U, class... A>
    void construct(U* p, A&&... a) {
        ::new (static_cast<void*>(p)) U(std::forward<A>(a)...);
    }
};

using BitEntryVec = std::vector<BitEntry, NoInitAlloc<BitEntry>>;

// The groups' entries are the concatenation of `chunks` (chunk t starts at
// entry chunk_off[t]; hdr[g].entry_off indexes the concatenation).  A
// caller that builds into the same tables again reuses their memory.
struct BitGroupTables {
    std::vector<BitGroupHdr> hdr;
    std::vector<int32_t> rows;        // kBitRowStride per group, -1 for unused
    std::vector<BitEntryVec> chunks;
    std::vector<int64_t> chunk_off;   // chunks.size() + 1 offsets
    std::vector<int32_t> row_group;   // group of each row, -1 for merge rows
    std::vector<int32_t> merge_rows;  // rows served by the merge kernel, ascending
    int64_t nentries() const { return chunk_off.empty() ? 0 : chunk_off.back(); }
    const BitEntry& entry(int64_t i) const {
        size_t t = 0;
        while (i >= chunk_off[t + 1]) ++t;
        return chunks[t][(size_t)(i - chunk_off[t])];
    }
};

// Consecutive rows form groups; with by_mag the eligible rows are first
// ordered by (magnitude, index) so that a group's rows have close
// magnitudes.  A group whose keys cannot be placed gives its last rows
// back to the next group until they can.
void build_bit_groups(const BitGroupInput& in, bool by_mag, BitGroupTables& out, int nthreads);

}  // namespace siml
