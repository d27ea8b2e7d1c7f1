// CPU replay of the bit-sliced row groups (test infrastructure).
//
// Builds the groups of a random reference set with build_bit_groups, streams
// every query molecule's levels with their slot words exactly as the column
// level stream holds them, probes the group tables as the kernels do (both
// candidate slots, key equality) and sums the row masks; each row's count
// must equal the multiset intersection computed by a sorted merge.
//
// usage: bitgroups_replay SEED NROWS NCOLS NONASCII_PER_MILLE BIG_PER_MILLE [PLANT]
//
// With PLANT = 1, three ASCII ids whose level-1 keys share both candidate
// slots are planted: a row holding two of them followed by a row holding
// the third (the group's cuckoo insertion runs out of displacements and
// the second row is given back to the next group), and a row holding all
// three (it cannot be placed at all and goes to the merge kernel).
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <unordered_map>
#include <vector>

#include "bitgroups.h"

using namespace siml;

struct Set {
    std::vector<int32_t> ids, counts, len;
    std::vector<int64_t> start;
};

static Set make_set(std::mt19937& rng, int n, int nonascii, int big) {
    Set s;
    s.start.push_back(0);
    for (int i = 0; i < n; ++i) {
        const int L = 1 + (int)(rng() % 120);
        std::map<int32_t, int32_t> m;
        for (int k = 0; k < L; ++k) {
            int32_t x = (int32_t)((rng() % 6) | ((rng() % 5) << 8) | ((rng() % 4) << 16) | ((rng() % 3) << 24));
            if ((int)(rng() % 1000) < nonascii) x |= (int32_t)0x80;
            m[x] += ((int)(rng() % 1000) < big) ? 17 : 1;
        }
        for (auto& kv : m) {
            s.ids.push_back(kv.first);
            s.counts.push_back(kv.second);
        }
        s.len.push_back((int32_t)m.size());
        s.start.push_back((int64_t)s.ids.size());
    }
    return s;
}

// Three ASCII ids whose level-1 keys have the same two candidate slots.
static std::vector<int32_t> colliding_ids() {
    std::unordered_map<uint32_t, std::vector<int32_t>> by_pair;
    for (int32_t b1 = 0x21; b1 < 0x7F; ++b1)
        for (int32_t b2 = 0x21; b2 < 0x7F; ++b2)
            for (int32_t b3 = 0x21; b3 < 0x7F; ++b3) {
                const int32_t x = 0x43 | (b1 << 8) | (b2 << 16) | (b3 << 24);
                uint32_t h1, h2;
                key_slots(level_key(x, 1u), h1, h2);
                std::vector<int32_t>& v = by_pair[std::min(h1, h2) * (uint32_t)kBitSlots + std::max(h1, h2)];
                v.push_back(x);
                if (v.size() == 3) return v;
            }
    return {};
}

static void add_row(Set& s, std::vector<int32_t> ids) {
    std::sort(ids.begin(), ids.end());
    for (int32_t x : ids) {
        s.ids.push_back(x);
        s.counts.push_back(1);
    }
    s.len.push_back((int32_t)ids.size());
    s.start.push_back((int64_t)s.ids.size());
}

int main(int argc, char** argv) {
    if (argc < 6) return 2;
    std::mt19937 rng((unsigned)atoi(argv[1]));
    int nr = atoi(argv[2]), nc = atoi(argv[3]);
    const int nonascii = atoi(argv[4]), big = atoi(argv[5]), plant = argc > 6 ? atoi(argv[6]) : 0;
    Set R = make_set(rng, nr, nonascii, big), C = make_set(rng, nc, nonascii, big);
    if (plant) {
        const std::vector<int32_t> x = colliding_ids();
        if (x.size() != 3) return 3;
        add_row(R, {x[0], x[1], 0x01010101});
        add_row(R, {x[2], 0x02010101});
        add_row(R, {x[0], x[1], x[2]});
        add_row(C, {x[0], x[1], x[2]});
        add_row(C, {x[2], 0x02010101});
        nr += 3;
        nc += 2;
    }
    std::vector<int32_t> mag(nr);
    for (int r = 0; r < nr; ++r)
        for (int64_t j = R.start[r]; j < R.start[r + 1]; ++j) mag[r] += R.counts[j];
    long bad = 0, checked = 0, grouped_rows = 0;
    for (int by_mag = 0; by_mag < 2; ++by_mag) {
        BitGroupInput in{R.ids.data(), R.counts.data(), R.start.data(), R.len.data(), mag.data(), nr};
        BitGroupTables bt;
        build_bit_groups(in, by_mag != 0, bt, 4);
        // every row is in exactly one group or in merge_rows
        std::vector<int> seen(nr, 0);
        for (size_t g = 0; g < bt.hdr.size(); ++g)
            for (int r = 0; r < bt.hdr[g].nrows; ++r) seen[bt.rows[g * kBitRowStride + r]]++;
        for (int32_t r : bt.merge_rows) seen[r]++;
        for (int r = 0; r < nr; ++r)
            if (seen[r] != 1) {
                printf("row %d covered %d times\n", r, seen[r]);
                return 1;
            }
        for (size_t g = 0; g < bt.hdr.size(); ++g) {
            const BitGroupHdr& h = bt.hdr[g];
            if (h.nrows < 1 || h.nrows > kBitRows || h.nkeys > kBitMaxKeys) return 1;
            grouped_rows += h.nrows;
            std::vector<uint32_t> key(kBitSlots, 0), mask(kBitSlots, 0);
            for (int k = 0; k < h.nkeys; ++k) {
                const BitEntry& e = bt.entry(h.entry_off + k);
                if (e.slot < 0 || e.slot >= kBitSlots || key[e.slot] != 0 || e.key == 0) return 1;
                uint32_t s1, s2;
                key_slots(e.key, s1, s2);
                if ((uint32_t)e.slot != s1 && (uint32_t)e.slot != s2) return 1;  // not a candidate slot
                key[e.slot] = e.key;
                mask[e.slot] = e.mask;
            }
            for (int c = 0; c < nc; ++c) {
                int cnt[32] = {0};
                for (int64_t j = C.start[c]; j < C.start[c + 1]; ++j) {
                    if (!ascii_id(C.ids[j])) continue;
                    for (int t = 1; t <= std::min(C.counts[j], kBitMaxLevel); ++t) {
                        const uint32_t k = level_key(C.ids[j], (uint32_t)t), sw = slot_word(k);
                        const uint32_t a1 = (sw & 0xFFFFu) / 8u, a2 = (sw >> 16) / 8u;
                        const uint32_t m = key[a1] == k ? mask[a1] : (key[a2] == k ? mask[a2] : 0u);
                        for (int r = 0; r < kBitRows; ++r) cnt[r] += (m >> r) & 1u;
                    }
                }
                for (int r = 0; r < h.nrows; ++r) {
                    const int row = bt.rows[g * kBitRowStride + r];
                    int isct = 0;
                    int64_t i = R.start[row], j = C.start[c];
                    while (i < R.start[row + 1] && j < C.start[c + 1]) {
                        if (R.ids[i] == C.ids[j]) {
                            isct += std::min(R.counts[i], C.counts[j]);
                            ++i;
                            ++j;
                        } else if (R.ids[i] < C.ids[j]) {
                            ++i;
                        } else {
                            ++j;
                        }
                    }
                    ++checked;
                    if (isct != cnt[r]) ++bad;
                }
            }
        }
        if (plant) {
            // the row with all three colliding keys cannot be placed
            if (std::find(bt.merge_rows.begin(), bt.merge_rows.end(), nr - 1) == bt.merge_rows.end()) {
                printf("row with three colliding keys is not a merge row\n");
                return 1;
            }
            if (bt.row_group[nr - 3] < 0 || bt.row_group[nr - 2] < 0 ||
                bt.row_group[nr - 3] == bt.row_group[nr - 2]) {
                printf("colliding rows share a group\n");
                return 1;
            }
        }
    }
    printf("checked %ld pairs over %ld grouped rows, %ld differ\n", checked, grouped_rows, bad);
    return bad ? 1 : 0;
}
