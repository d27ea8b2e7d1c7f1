// bitgroups.cpp -- host builder of the bit-sliced row groups (bitgroups.h).
#include "bitgroups.h"

#include <algorithm>
#include <thread>

namespace siml {

namespace {

// Levels of one row as sorted 64-bit keys (y << 4 | t).
bool row_levels(const BitGroupInput& in, int64_t r, std::vector<uint64_t>& out) {
    out.clear();
    const int32_t L = in.len[r];
    const int32_t* id = in.ids + in.start[r];
    const int32_t* ct = in.counts + in.start[r];
    for (int32_t k = 0; k < L; ++k) {
        if (ct[k] < 0 || ct[k] > kBitMaxLevel) return false;
        if (k && id[k] <= id[k - 1]) return false;
        for (int32_t t = 1; t <= ct[k]; ++t) out.push_back(((uint64_t)level_key(id[k], (uint32_t)t) << 4) | (uint64_t)t);
    }
    if ((int64_t)out.size() > kBitMaxKeys) return false;
    std::sort(out.begin(), out.end());
    return true;
}

int64_t union_count(const std::vector<uint64_t>& u, const std::vector<uint64_t>& v) {
    size_t i = 0, j = 0;
    int64_t n = 0;
    while (i < u.size() && j < v.size()) {
        if (u[i] == v[j]) {
            ++i;
            ++j;
        } else if (u[i] < v[j]) {
            ++i;
        } else {
            ++j;
        }
        ++n;
    }
    return n + (int64_t)(u.size() - i) + (int64_t)(v.size() - j);
}

inline uint64_t next_rand(uint64_t& s) {
    s ^= s >> 12;
    s ^= s << 25;
    s ^= s >> 27;
    return s * 2685821657736338717ull;
}

bool place(const std::vector<uint32_t>& ys, const BitGroupHdr& h, std::vector<int32_t>& tab) {
    tab.assign(kBitSlots, -1);
    const int K = (int)ys.size();
    auto slot = [&](int k, int which) {
        return which ? bit_hash(ys[(size_t)k], h.a2, h.b2) : bit_hash(ys[(size_t)k], h.a1, h.b1);
    };
    for (int k = 0; k < K; ++k) {
        int cur = k;
        uint32_t pos = slot(cur, 0);
        bool placed = false;
        for (int it = 0; it < 4 * kBitSlots; ++it) {
            if (tab[pos] < 0) {
                tab[pos] = cur;
                placed = true;
                break;
            }
            std::swap(cur, tab[pos]);
            const uint32_t p1 = slot(cur, 0), p2 = slot(cur, 1);
            const uint32_t np = (pos == p1) ? p2 : p1;
            if (np == pos) return false;
            pos = np;
        }
        if (!placed) return false;
    }
    return true;
}

}  // namespace

void build_bit_groups(const BitGroupInput& in, bool by_mag, BitGroupTables& out, int nthreads) {
    out = BitGroupTables();
    out.row_group.assign((size_t)in.n, -1);
    // 1. eligible rows, in the grouping order
    std::vector<int32_t> order;
    std::vector<uint64_t> lv;
    for (int64_t r = 0; r < in.n; ++r) {
        if (row_levels(in, r, lv))
            order.push_back((int32_t)r);
        else
            out.merge_rows.push_back((int32_t)r);
    }
    if (by_mag)
        std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return in.mag[a] < in.mag[b]; });
    // 2. greedy grouping
    std::vector<std::vector<int32_t>> plans;
    std::vector<uint64_t> u, tmp;
    std::vector<int32_t> cur;
    for (int32_t r : order) {
        row_levels(in, r, lv);
        if (!cur.empty() && ((int)cur.size() == kBitRows || union_count(u, lv) > kBitMaxKeys)) {
            plans.push_back(cur);
            cur.clear();
            u.clear();
        }
        tmp.resize(u.size() + lv.size());
        tmp.resize((size_t)(std::set_union(u.begin(), u.end(), lv.begin(), lv.end(), tmp.begin()) - tmp.begin()));
        u.swap(tmp);
        cur.push_back(r);
    }
    if (!cur.empty()) plans.push_back(cur);
    const int64_t G = (int64_t)plans.size();
    std::vector<int64_t> nk((size_t)G);
    std::vector<char> failed((size_t)std::max<int64_t>(G, 1), 0);
    std::vector<std::vector<BitEntry>> ents((size_t)G);
    out.hdr.resize((size_t)G);
    // 3. per group: union with masks, cuckoo placement (parallel)
    auto work = [&](int64_t g0, int64_t g1) {
        std::vector<uint64_t> lvr;
        std::vector<std::pair<uint64_t, uint32_t>> kv;
        std::vector<uint32_t> ys;
        std::vector<int32_t> tab;
        for (int64_t g = g0; g < g1; ++g) {
            const std::vector<int32_t>& rows = plans[(size_t)g];
            kv.clear();
            int32_t maxsum = 0;
            for (size_t i = 0; i < rows.size(); ++i) {
                row_levels(in, rows[i], lvr);
                for (uint64_t k : lvr) kv.emplace_back(k, 1u << i);
                int32_t s = 0;
                const int32_t* ct = in.counts + in.start[rows[i]];
                for (int32_t k = 0; k < in.len[rows[i]]; ++k) s += ct[k];
                maxsum = std::max(maxsum, s);
            }
            std::sort(kv.begin(), kv.end());
            std::vector<std::pair<uint64_t, uint32_t>> keys;
            for (auto& e : kv) {
                if (!keys.empty() && keys.back().first == e.first)
                    keys.back().second |= e.second;
                else
                    keys.push_back(e);
            }
            ys.resize(keys.size());
            for (size_t k = 0; k < keys.size(); ++k) ys[k] = (uint32_t)(keys[k].first >> 4);
            BitGroupHdr& h = out.hdr[(size_t)g];
            uint64_t rs = 0x243F6A8885A308D3ull ^ (uint64_t)(g + 1) * 0x9E3779B97F4A7C15ull;
            bool ok = false;
            for (int attempt = 0; attempt < 64 && !ok; ++attempt) {
                h.a1 = (uint32_t)(next_rand(rs) >> 40) | 1u;
                h.b1 = (uint32_t)(next_rand(rs) >> 40) | 1u;
                h.a2 = (uint32_t)(next_rand(rs) >> 40) | 1u;
                h.b2 = (uint32_t)(next_rand(rs) >> 40) | 1u;
                if (h.a1 == h.a2 && h.b1 == h.b2) continue;
                ok = place(ys, h, tab);
            }
            if (!ok) {
                failed[(size_t)g] = 1;
                continue;
            }
            h.nrows = (int32_t)rows.size();
            h.nkeys = (int32_t)keys.size();
            h.nbits = maxsum <= 255 ? 8 : 16;
            h.pad0 = 0;
            std::vector<BitEntry>& E = ents[(size_t)g];
            E.clear();
            for (int s = 0; s < kBitSlots; ++s) {
                const int k = tab[(size_t)s];
                if (k < 0) continue;
                const uint32_t t = (uint32_t)(keys[(size_t)k].first & 15u);
                E.push_back(BitEntry{ys[(size_t)k], keys[(size_t)k].second | (t << 28), s, 0});
            }
            nk[(size_t)g] = (int64_t)E.size();
        }
    };
    unsigned hw = std::thread::hardware_concurrency();
    const int nt = nthreads > 0 ? nthreads : (int)std::min<unsigned>(hw ? hw : 1, 32);
    if (G < 64 || nt <= 1) {
        work(0, G);
    } else {
        std::vector<std::thread> th;
        const int64_t chunk = (G + nt - 1) / nt;
        for (int t = 0; t < nt; ++t) {
            const int64_t a = t * chunk, b = std::min(G, a + chunk);
            if (a >= b) break;
            th.emplace_back(work, a, b);
        }
        for (auto& t : th) t.join();
    }
    // 4. gather successful groups; failed groups' rows go to the merge path
    std::vector<BitGroupHdr> hdr;
    for (int64_t g = 0; g < G; ++g) {
        const std::vector<int32_t>& rows = plans[(size_t)g];
        if (failed[(size_t)g]) {
            out.merge_rows.insert(out.merge_rows.end(), rows.begin(), rows.end());
            continue;
        }
        BitGroupHdr h = out.hdr[(size_t)g];
        h.entry_off = (int64_t)out.entries.size();
        out.entries.insert(out.entries.end(), ents[(size_t)g].begin(), ents[(size_t)g].end());
        const int32_t gi = (int32_t)hdr.size();
        for (int i = 0; i < kBitRowStride; ++i) {
            const int32_t r = i < (int)rows.size() ? rows[(size_t)i] : -1;
            out.rows.push_back(r);
            if (r >= 0) out.row_group[(size_t)r] = gi;
        }
        hdr.push_back(h);
    }
    out.hdr.swap(hdr);
    std::sort(out.merge_rows.begin(), out.merge_rows.end());
}

}  // namespace siml
