// bitgroups.h -- row groups for the bit-sliced multi-row kernels.
//
// Counting levels.  For counts ca, cb >= 0,
//     min(ca, cb) = #{ t >= 1 : t <= ca and t <= cb },
// so the intersection sum_x min(ca(x), cb(x)) of two Lingo multisets is the
// size of the intersection of their LEVEL sets {(x, t) : 1 <= t <= c(x)}.
// A level (x, t) is encoded as the pair
//     y = x ^ ((t - 1) * 0x9E3779B9),  w = t << 28            (t in 1..15)
// which is injective for fixed t.  The query (column) set is streamed as
// these (y, w) pairs, levels above 15 dropped: no row of a group has them.
//
// A group is up to kBitRows reference rows whose level union has at most
// kBitMaxKeys levels.  Each level of the union occupies one slot of a
// two-choice cuckoo table of kBitSlots slots, at h1(y) or h2(y) with
//     h_i(y) = ((umul24(y, a_i) + umul24(y >> 8, b_i)) mod 2^32) >> (32 - kBitSlotBits),
// holding the pair (y, mask | w), where bit r of mask (r < 28) says that row
// r of the group has the level.  A probe with (y, w) matches a slot iff the
// ids are equal and the top 4 bits equal t; empty slots are (0, 0), which no
// probe (t >= 1) matches.  Summing the matched masks as 28 vertical binary
// counters gives every row's intersection with the column at once.
//
// Rows qualify when their ids are strictly ascending, every count is in
// [0, 15] and the row alone fits in a group; other rows use the merge
// kernel.
#pragma once

#include <cstdint>
#include <vector>

namespace siml {

constexpr int kBitRows = 28;
constexpr int kBitRowStride = 32;  // row indices stored per group
constexpr int kBitSlotBits = 12;
constexpr int kBitSlots = 1 << kBitSlotBits;  // 4096 slots x 8 B = 32 KiB of LDS
constexpr int kBitMaxKeys = 1800;             // load factor <= 0.44
constexpr int kBitMaxLevel = 15;
constexpr uint32_t kLevelMul = 0x9E3779B9u;

inline uint32_t level_key(int32_t x, uint32_t t) { return (uint32_t)x ^ ((t - 1u) * kLevelMul); }

struct BitGroupHdr {
    int32_t nrows;      // 1..kBitRows
    int32_t nkeys;      // levels in the union
    int32_t nbits;      // bit planes the counts need: 8 or 16
    int32_t pad0;
    uint32_t a1, b1;    // multipliers of h1 (24-bit)
    uint32_t a2, b2;    // multipliers of h2 (24-bit)
    int64_t entry_off;  // first entry of the group in the entry arrays
};

struct BitEntry {
    uint32_t y;     // level key
    uint32_t word;  // row mask | t << 28
    int32_t slot;
    int32_t pad;
};

inline uint32_t umul24_host(uint32_t u, uint32_t v) { return (uint32_t)((uint64_t)(u & 0xFFFFFFu) * (v & 0xFFFFFFu)); }
inline uint32_t bit_hash(uint32_t y, uint32_t a, uint32_t b) {
    return (umul24_host(y, a) + umul24_host(y >> 8, b)) >> (32 - kBitSlotBits);
}

struct BitGroupInput {
    const int32_t* ids;
    const int32_t* counts;
    const int64_t* start;
    const int32_t* len;
    const int32_t* mag;  // groups are formed over rows sorted by magnitude when by_mag
    int64_t n;
};

struct BitGroupTables {
    std::vector<BitGroupHdr> hdr;
    std::vector<int32_t> rows;        // kBitRowStride per group, -1 for unused
    std::vector<BitEntry> entries;
    std::vector<int32_t> row_group;   // group of each row, -1 for merge rows
    std::vector<int32_t> merge_rows;  // rows served by the merge kernel, ascending
};

// Consecutive rows form groups; with by_mag the eligible rows are first
// ordered by (magnitude, index) so that a group's rows have close
// magnitudes.
void build_bit_groups(const BitGroupInput& in, bool by_mag, BitGroupTables& out, int nthreads);

}  // namespace siml
