// bitgroups.cpp -- host builder of the bit-sliced row groups (bitgroups.h).
#include "bitgroups.h"
#include "host_threads.h"

#include <algorithm>
#include <cstdlib>
#include <thread>

namespace siml {

namespace {

// Whether row r qualifies for a group: ids strictly ascending and ASCII,
// counts in [0, kBitMaxLevel], at most kBitMaxKeys levels.
bool row_qualifies(const BitGroupInput& in, int64_t r) {
    const int32_t L = in.len[r];
    const int32_t* id = in.ids + in.start[r];
    const int32_t* ct = in.counts + in.start[r];
    int64_t nlev = 0;
    for (int32_t k = 0; k < L; ++k) {
        if (ct[k] < 0 || ct[k] > kBitMaxLevel || !ascii_id(id[k])) return false;
        if (k && id[k] <= id[k - 1]) return false;
        nlev += ct[k];
    }
    return nlev <= kBitMaxKeys;
}

// Level keys of a qualifying row, in id order (distinct).
void row_levels(const BitGroupInput& in, int64_t r, std::vector<uint32_t>& out) {
    out.clear();
    const int32_t L = in.len[r];
    const int32_t* id = in.ids + in.start[r];
    const int32_t* ct = in.counts + in.start[r];
    for (int32_t k = 0; k < L; ++k)
        for (int32_t t = 1; t <= ct[k]; ++t) out.push_back(level_key(id[k], (uint32_t)t));
}

// A group's table while its rows are added: key and row mask per slot,
// key 0 for an empty slot (no level key is 0).
struct GroupTable {
    uint32_t key[kBitSlots];
    uint32_t mask[kBitSlots];
    std::vector<uint32_t> used;  // slots filled since the last clear
    GroupTable() {
        std::fill(key, key + kBitSlots, 0u);
        std::fill(mask, mask + kBitSlots, 0u);
    }
    void clear() {
        for (uint32_t s : used) key[s] = mask[s] = 0u;
        used.clear();
    }
    int find(uint32_t k, uint32_t h1, uint32_t h2) const {
        if (key[h1] == k) return (int)h1;
        if (key[h2] == k) return (int)h2;
        return -1;
    }
    // two-choice cuckoo insertion of a key that is not in the table; false
    // (a key lost from the table) after 4 * kBitSlots displacements
    bool insert(uint32_t k, uint32_t m, uint32_t h1, uint32_t h2) {
        uint32_t pos = key[h1] == 0u ? h1 : h2;
        for (int it = 0; it < 4 * kBitSlots; ++it) {
            if (key[pos] == 0u) {
                key[pos] = k;
                mask[pos] = m;
                used.push_back(pos);
                return true;
            }
            std::swap(k, key[pos]);
            std::swap(m, mask[pos]);
            uint32_t a, b;
            key_slots(k, a, b);
            pos = pos == a ? b : a;
        }
        return false;
    }
};

// Two-choice cuckoo placement of `keys` with the fixed slot functions;
// tab[s] = index of the key in slot s or -1.
bool place(const std::vector<uint32_t>& keys, std::vector<int32_t>& tab) {
    tab.assign(kBitSlots, -1);
    const int K = (int)keys.size();
    std::vector<uint32_t> s1((size_t)K), s2((size_t)K);
    for (int k = 0; k < K; ++k) key_slots(keys[(size_t)k], s1[(size_t)k], s2[(size_t)k]);
    for (int k = 0; k < K; ++k) {
        int cur = k;
        uint32_t pos = s1[(size_t)cur];
        bool placed = false;
        for (int it = 0; it < 4 * kBitSlots; ++it) {
            if (tab[pos] < 0) {
                tab[pos] = cur;
                placed = true;
                break;
            }
            std::swap(cur, tab[pos]);
            pos = (pos == s1[(size_t)cur]) ? s2[(size_t)cur] : s1[(size_t)cur];
        }
        if (!placed) return false;
    }
    return true;
}

// A group of a chunk: its rows and header (entry_off relative to the
// chunk's entry array until the gather).
struct Built {
    int32_t rows[kBitRowStride];
    BitGroupHdr hdr;
};

// Header and slot-ordered entries of a placed group, appended to the
// chunk's groups and entries.
void emit_group(const std::vector<int32_t>& rows, const std::vector<uint32_t>& keys,
                const std::vector<uint32_t>& masks, const std::vector<int32_t>& tab, int32_t maxsum,
                std::vector<Built>& out, BitEntryVec& ents) {
    Built b;
    for (int i = 0; i < kBitRowStride; ++i) b.rows[i] = i < (int)rows.size() ? rows[(size_t)i] : -1;
    b.hdr.nrows = (int32_t)rows.size();
    b.hdr.nkeys = (int32_t)keys.size();
    b.hdr.nbits = maxsum <= 255 ? 8 : 16;
    b.hdr.pad0 = 0;
    b.hdr.entry_off = (int64_t)ents.size();
    for (int s = 0; s < kBitSlots; ++s) {
        const int k = tab[(size_t)s];
        if (k >= 0) ents.push_back(BitEntry{keys[(size_t)k], masks[(size_t)k], s, 0});
    }
    out.push_back(b);
}

// Places the sorted key union u of `rows` (level lists lvs[i]) and emits the
// group; false (nothing emitted) when the keys cannot be placed.  Each level
// finds its key's index through the placed table (slot h1, else h2) and
// sets its row's mask bit.
bool try_group(const std::vector<int32_t>& rows, const std::vector<std::vector<uint32_t>>& lvs,
               const std::vector<uint32_t>& u, std::vector<int32_t>& tab, std::vector<uint32_t>& masks,
               std::vector<Built>& out, BitEntryVec& ents) {
    if (!place(u, tab)) return false;
    masks.assign(u.size(), 0u);
    int32_t maxsum = 0;
    for (size_t i = 0; i < rows.size(); ++i) {
        const std::vector<uint32_t>& lv = lvs[i];
        for (uint32_t k : lv) {
            uint32_t h1, h2;
            key_slots(k, h1, h2);
            int32_t j = tab[h1];
            if (j < 0 || u[(size_t)j] != k) j = tab[h2];
            masks[(size_t)j] |= 1u << i;
        }
        maxsum = std::max(maxsum, (int32_t)lv.size());
    }
    emit_group(rows, u, masks, tab, maxsum, out, ents);
    return true;
}

}  // namespace

void build_bit_groups(const BitGroupInput& in, bool by_mag, BitGroupTables& out, int nthreads) {
    out.hdr.clear();
    out.rows.clear();
    out.row_group.assign((size_t)in.n, -1);
    out.merge_rows.clear();
    const int nt = nthreads > 0 ? nthreads : default_threads(32);
    auto run = [&](int64_t n, auto&& fn) {
        if (n < 256 || nt <= 1) {
            fn(0, n, 0);
            return;
        }
        std::vector<std::thread> th;
        const int64_t chunk = (n + nt - 1) / nt;
        for (int t = 0; t < nt; ++t) {
            const int64_t a = t * chunk, b = std::min(n, a + chunk);
            if (a >= b) break;
            th.emplace_back(fn, a, b, t);
        }
        for (auto& t : th) t.join();
    };
    // 1. eligibility
    std::vector<char> ok((size_t)in.n, 0);
    run(in.n, [&](int64_t a, int64_t b, int) {
        for (int64_t r = a; r < b; ++r) ok[(size_t)r] = row_qualifies(in, r) ? 1 : 0;
    });
    std::vector<int32_t> order;
    for (int64_t r = 0; r < in.n; ++r) {
        if (ok[(size_t)r])
            order.push_back((int32_t)r);
        else
            out.merge_rows.push_back((int32_t)r);
    }
    if (by_mag)
        std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return in.mag[a] < in.mag[b]; });
    // 2. greedy grouping of consecutive rows, each group's table built as
    //    soon as it closes.  Chunks of at least kChunkRows rows are grouped
    //    independently (in parallel); each chunk ends with a partial group,
    //    so chunks are kept long enough for that to cost at most a group in
    //    64.
    const int64_t m = (int64_t)order.size();
    constexpr int64_t kChunkRows = 64 * kBitRows;
    const int nchunks = (int)std::max<int64_t>(1, std::min<int64_t>(nt, m / kChunkRows));
    std::vector<std::vector<Built>> built((size_t)nchunks);
    out.chunks.resize((size_t)nchunks);
    for (BitEntryVec& c : out.chunks) c.clear();
    std::vector<std::vector<int32_t>> merged((size_t)nchunks);
    auto greedy = [&](int64_t a, int64_t b, int t) {
        std::unique_ptr<GroupTable> T(new GroupTable);
        std::vector<uint32_t> u, lv;
        std::vector<std::vector<uint32_t>> lvs((size_t)kBitRows);
        std::vector<int32_t> cur, tab;
        std::vector<uint32_t> masks;
        int64_t i = a;
        while (i < b) {
            // add rows to the table while the group has room and its key
            // union stays within kBitMaxKeys
            T->clear();
            cur.clear();
            int nkeys = 0, maxsum = 0;
            bool lost = false;
            int64_t j = i;
            for (; j < b && (int)cur.size() < kBitRows && !lost; ++j) {
                row_levels(in, order[(size_t)j], lv);
                const uint32_t bit = 1u << cur.size();
                int nnew = 0;
                for (size_t q = 0; q < lv.size() && !lost; ++q) {
                    uint32_t h1, h2;
                    key_slots(lv[q], h1, h2);
                    const int sl = T->find(lv[q], h1, h2);
                    if (sl >= 0) {
                        T->mask[sl] |= bit;
                    } else {
                        ++nnew;
                        lost = !T->insert(lv[q], bit, h1, h2);
                    }
                }
                if (!lost && nkeys + nnew > kBitMaxKeys) {
                    // the row does not fit: take its levels out again (every
                    // other key stays in one of its two slots)
                    for (uint32_t k : lv) {
                        uint32_t h1, h2;
                        key_slots(k, h1, h2);
                        const int sl = T->find(k, h1, h2);
                        T->mask[sl] &= ~bit;
                        if (T->mask[sl] == 0u) T->key[sl] = 0u;
                    }
                    break;
                }
                nkeys += nnew;
                maxsum = std::max(maxsum, (int)lv.size());
                lvs[cur.size()].swap(lv);
                cur.push_back(order[(size_t)j]);
            }
            if (!lost) {
                BitEntryVec& ents = out.chunks[(size_t)t];
                Built g;
                for (int r = 0; r < kBitRowStride; ++r) g.rows[r] = r < (int)cur.size() ? cur[(size_t)r] : -1;
                g.hdr.nrows = (int32_t)cur.size();
                g.hdr.nkeys = nkeys;
                g.hdr.nbits = maxsum <= 255 ? 8 : 16;
                g.hdr.pad0 = 0;
                g.hdr.entry_off = (int64_t)ents.size();
                for (int sl = 0; sl < kBitSlots; ++sl)
                    if (T->key[sl]) ents.push_back(BitEntry{T->key[sl], T->mask[sl], sl, 0});
                built[(size_t)t].push_back(g);
                i = j;
                continue;
            }
            // an insertion ran out of displacements: place the sorted union
            // afresh, giving the group's last rows back to the next group
            // until the keys fit; a single row that does not fit goes to the
            // merge kernel
            for (;;) {
                u.clear();
                for (size_t r = 0; r < cur.size(); ++r) u.insert(u.end(), lvs[r].begin(), lvs[r].end());
                std::sort(u.begin(), u.end());
                u.erase(std::unique(u.begin(), u.end()), u.end());
                if (try_group(cur, lvs, u, tab, masks, built[(size_t)t], out.chunks[(size_t)t])) break;
                if (cur.size() == 1) {
                    merged[(size_t)t].push_back(cur[0]);
                    break;
                }
                cur.pop_back();
                --j;
            }
            i = j;
        }
    };
    if (nchunks == 1) {
        greedy(0, m, 0);
    } else {
        std::vector<std::thread> th;
        for (int t = 0; t < nchunks; ++t) th.emplace_back(greedy, m * t / nchunks, m * (t + 1) / nchunks, t);
        for (auto& x : th) x.join();
    }
    // 3. gather the headers and rows; entry offsets become offsets into the
    //    concatenation of the chunks' entry arrays
    int64_t ng = 0;
    out.chunk_off.assign((size_t)nchunks + 1, 0);
    for (int t = 0; t < nchunks; ++t) {
        ng += (int64_t)built[(size_t)t].size();
        out.chunk_off[(size_t)t + 1] = out.chunk_off[(size_t)t] + (int64_t)out.chunks[(size_t)t].size();
        out.merge_rows.insert(out.merge_rows.end(), merged[(size_t)t].begin(), merged[(size_t)t].end());
    }
    out.hdr.resize((size_t)ng);
    out.rows.resize((size_t)ng * kBitRowStride);
    int64_t gi = 0;
    for (int t = 0; t < nchunks; ++t)
        for (const Built& b : built[(size_t)t]) {
            BitGroupHdr h = b.hdr;
            h.entry_off += out.chunk_off[(size_t)t];
            out.hdr[(size_t)gi] = h;
            for (int i = 0; i < kBitRowStride; ++i) {
                const int32_t r = b.rows[i];
                out.rows[(size_t)(gi * kBitRowStride + i)] = r;
                if (r >= 0) out.row_group[(size_t)r] = (int32_t)gi;
            }
            ++gi;
        }
    std::sort(out.merge_rows.begin(), out.merge_rows.end());
}

}  // namespace siml
