// bitgroups.cpp -- host builder of the bit-sliced row groups (bitgroups.h).
#include "bitgroups.h"
#include "host_threads.h"

#include <algorithm>
#include <cstdlib>
#include <thread>

namespace siml {

namespace {

// Level keys of one row, sorted; false when the row does not qualify.
bool row_levels(const BitGroupInput& in, int64_t r, std::vector<uint32_t>& out) {
    out.clear();
    const int32_t L = in.len[r];
    const int32_t* id = in.ids + in.start[r];
    const int32_t* ct = in.counts + in.start[r];
    for (int32_t k = 0; k < L; ++k) {
        if (ct[k] < 0 || ct[k] > kBitMaxLevel || !ascii_id(id[k])) return false;
        if (k && id[k] <= id[k - 1]) return false;
        for (int32_t t = 1; t <= ct[k]; ++t) out.push_back(level_key(id[k], (uint32_t)t));
        if ((int64_t)out.size() > kBitMaxKeys) return false;
    }
    std::sort(out.begin(), out.end());
    return true;
}

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

struct Built {
    std::vector<int32_t> rows;
    BitGroupHdr hdr;
    std::vector<BitEntry> ents;
};

// Header and slot-ordered entries of a placed group.
void emit_group(const std::vector<int32_t>& rows, const std::vector<uint32_t>& keys,
                const std::vector<uint32_t>& masks, const std::vector<int32_t>& tab, int32_t maxsum,
                std::vector<Built>& out) {
    Built b;
    b.rows = rows;
    b.hdr.nrows = (int32_t)rows.size();
    b.hdr.nkeys = (int32_t)keys.size();
    b.hdr.nbits = maxsum <= 255 ? 8 : 16;
    b.hdr.pad0 = 0;
    b.hdr.entry_off = 0;
    b.ents.reserve(keys.size());
    for (int s = 0; s < kBitSlots; ++s) {
        const int k = tab[(size_t)s];
        if (k >= 0) b.ents.push_back(BitEntry{keys[(size_t)k], masks[(size_t)k], s, 0});
    }
    out.push_back(std::move(b));
}

// Places the sorted key union u of `rows` (level lists lvs[i]) and emits the
// group; false (nothing emitted) when the keys cannot be placed.  Each level
// finds its key's index through the placed table (slot h1, else h2) and
// sets its row's mask bit.
bool try_group(const std::vector<int32_t>& rows, const std::vector<std::vector<uint32_t>>& lvs,
               const std::vector<uint32_t>& u, std::vector<int32_t>& tab, std::vector<uint32_t>& masks,
               std::vector<Built>& out) {
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
    emit_group(rows, u, masks, tab, maxsum, out);
    return true;
}

}  // namespace

void build_bit_groups(const BitGroupInput& in, bool by_mag, BitGroupTables& out, int nthreads) {
    out = BitGroupTables();
    out.row_group.assign((size_t)in.n, -1);
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
        std::vector<uint32_t> lv;
        for (int64_t r = a; r < b; ++r) ok[(size_t)r] = row_levels(in, r, lv) ? 1 : 0;
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
    std::vector<std::vector<int32_t>> merged((size_t)nchunks);
    auto greedy = [&](int64_t a, int64_t b, int t) {
        std::vector<uint32_t> u, tmp, lv;
        std::vector<std::vector<uint32_t>> lvs((size_t)kBitRows);
        std::vector<int32_t> cur, tab;
        std::vector<uint32_t> masks;
        int64_t i = a;
        while (i < b) {
            // take rows while the group has room and its key union stays
            // within kBitMaxKeys
            cur.clear();
            int64_t j = i;
            for (; j < b && (int)cur.size() < kBitRows; ++j) {
                row_levels(in, order[(size_t)j], lv);
                if (cur.empty()) {
                    u.assign(lv.begin(), lv.end());
                } else {
                    tmp.resize(u.size() + lv.size());
                    tmp.resize((size_t)(std::set_union(u.begin(), u.end(), lv.begin(), lv.end(), tmp.begin()) -
                                        tmp.begin()));
                    if ((int64_t)tmp.size() > kBitMaxKeys) break;
                    u.swap(tmp);
                }
                lvs[cur.size()].swap(lv);
                cur.push_back(order[(size_t)j]);
            }
            // a failed placement gives the group's last row back to the next
            // group until the keys fit; a single row that does not fit goes
            // to the merge kernel
            while (!try_group(cur, lvs, u, tab, masks, built[(size_t)t])) {
                if (cur.size() == 1) {
                    merged[(size_t)t].push_back(cur[0]);
                    break;
                }
                cur.pop_back();
                --j;
                u.clear();
                for (size_t r = 0; r < cur.size(); ++r) u.insert(u.end(), lvs[r].begin(), lvs[r].end());
                std::sort(u.begin(), u.end());
                u.erase(std::unique(u.begin(), u.end()), u.end());
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
    // 3. gather: offsets first, then each chunk copies its groups in place
    std::vector<int64_t> g_off((size_t)nchunks + 1, 0), e_off((size_t)nchunks + 1, 0);
    for (int t = 0; t < nchunks; ++t) {
        int64_t ne = 0;
        for (const Built& b : built[(size_t)t]) ne += (int64_t)b.ents.size();
        g_off[(size_t)t + 1] = g_off[(size_t)t] + (int64_t)built[(size_t)t].size();
        e_off[(size_t)t + 1] = e_off[(size_t)t] + ne;
        out.merge_rows.insert(out.merge_rows.end(), merged[(size_t)t].begin(), merged[(size_t)t].end());
    }
    out.hdr.resize((size_t)g_off[(size_t)nchunks]);
    out.rows.resize((size_t)g_off[(size_t)nchunks] * kBitRowStride);
    out.entries.resize((size_t)e_off[(size_t)nchunks]);
    auto gather = [&](int t) {
        int64_t gi = g_off[(size_t)t], eo = e_off[(size_t)t];
        for (Built& b : built[(size_t)t]) {
            b.hdr.entry_off = eo;
            std::copy(b.ents.begin(), b.ents.end(), out.entries.begin() + eo);
            eo += (int64_t)b.ents.size();
            for (int i = 0; i < kBitRowStride; ++i) {
                const int32_t r = i < (int)b.rows.size() ? b.rows[(size_t)i] : -1;
                out.rows[(size_t)(gi * kBitRowStride + i)] = r;
                if (r >= 0) out.row_group[(size_t)r] = (int32_t)gi;
            }
            out.hdr[(size_t)gi] = b.hdr;
            ++gi;
            std::vector<BitEntry>().swap(b.ents);
        }
    };
    if (nchunks == 1) {
        gather(0);
    } else {
        std::vector<std::thread> th;
        for (int t = 0; t < nchunks; ++t) th.emplace_back(gather, t);
        for (auto& x : th) x.join();
    }
    std::sort(out.merge_rows.begin(), out.merge_rows.end());
}

}  // namespace siml
