// groups.h -- row groups for the hashed multi-row kernels.
//
// A group is up to kGroupRows consecutive reference rows.  The union of
// their Lingo ids (at most kMaxKeys) is placed in a two-choice cuckoo table
// of kSlots slots: id x lives in slot h1(x) or h2(x), where
//   h_i(x) = ((umul24(x, a_i) + umul24(x >> 8, b_i)) mod 2^32) >> (32 - kSlotBits)
// and umul24(u, v) is the low 32 bits of (u mod 2^24) * (v mod 2^24).  Both
// hashes see all 32 bits of x (x mod 2^24 and x >> 8 determine x).
// Slot 0 is never used, so a probe that matches neither slot reads the
// all-zero count vector of slot 0.  Each slot carries the id's count in each
// row of the group as eight u16 (zero where the row lacks the id).  Empty
// slots hold `empty_key`, an int32 that is not in the union.
//
// For sorted-unique rows and columns the intersection of row r with a
// column is then sum_j min(count_r(x_j), cb_j) over the column's entries,
// which is exactly the sum of min counts over the common ids computed by
// the sorted merge of _CPULingo.c:59-69.
#pragma once

#include <cstdint>
#include <vector>

namespace siml {

constexpr int kGroupRows = 8;
constexpr int kSlotBits = 11;
constexpr int kSlots = 1 << kSlotBits;  // 2048 slots: 8 KiB of ids + 32 KiB of counts in LDS
constexpr int kMaxKeys = 880;           // load factor <= 0.43 for the two-choice table
constexpr uint32_t kCountMax = 0xFFFFu;

struct GroupHdr {
    int32_t row0;       // first reference row of the group
    int32_t nrows;      // 1..kGroupRows
    int32_t nkeys;      // union size
    int32_t empty_key;  // id stored in empty slots (not in the union)
    uint32_t a1, b1;    // multipliers of h1 (24-bit)
    uint32_t a2, b2;    // multipliers of h2 (24-bit)
    int32_t pad;
    int64_t entry_off;  // first entry of the group in the entry arrays
};

struct GroupEntry {
    int32_t key;
    int32_t slot;
};

inline uint32_t umul24(uint32_t u, uint32_t v) { return (uint32_t)((uint64_t)(u & 0xFFFFFFu) * (v & 0xFFFFFFu)); }
inline uint32_t group_hash(uint32_t x, uint32_t a, uint32_t b) {
    return (umul24(x, a) + umul24(x >> 8, b)) >> (32 - kSlotBits);
}

// Host description of one reference set for grouping: molecule r has len[r]
// (id, count) pairs at ids/counts + start[r].
struct GroupInput {
    const int32_t* ids;
    const int32_t* counts;
    const int64_t* start;
    const int32_t* len;
    const int32_t* csum;
    int64_t n;
};

struct GroupTables {
    std::vector<GroupHdr> hdr;
    std::vector<GroupEntry> ekey;
    std::vector<uint32_t> ecnt;       // 4 words (8 x u16) per entry
    std::vector<int32_t> row_group;   // group of each row, -1 for rows on the merge path
    std::vector<int32_t> merge_rows;  // rows that cannot use the hashed kernels
};

// Rows qualify when their ids are strictly ascending, their counts lie in
// [0, 65535], their count sum is <= 65535 and they have <= kMaxKeys ids.
// Returns false only on an internal failure.
bool build_groups(const GroupInput& in, GroupTables& out, int nthreads);

}  // namespace siml
