// CPU replay of the bit-sliced row groups (test infrastructure).
//
// Builds the groups of a random reference set with build_bit_groups, streams
// every query molecule's levels with their slot words exactly as the column
// level stream holds them, probes the group tables as the kernels do (both
// candidate slots, key equality) and sums the row masks; each row's count
// must equal the multiset intersection computed by a sorted merge.
//
// usage: bitgroups_replay SEED NROWS NCOLS NONASCII_PER_MILLE BIG_PER_MILLE
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
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

int main(int argc, char** argv) {
    if (argc < 6) return 2;
    std::mt19937 rng((unsigned)atoi(argv[1]));
    const int nr = atoi(argv[2]), nc = atoi(argv[3]), nonascii = atoi(argv[4]), big = atoi(argv[5]);
    Set R = make_set(rng, nr, nonascii, big), C = make_set(rng, nc, nonascii, big);
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
                const BitEntry& e = bt.entries[(size_t)(h.entry_off + k)];
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
    }
    printf("checked %ld pairs over %ld grouped rows, %ld differ\n", checked, grouped_rows, bad);
    return bad ? 1 : 0;
}
