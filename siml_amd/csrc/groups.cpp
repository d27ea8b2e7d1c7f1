// groups.cpp -- host builder of the row-group cuckoo tables (see groups.h).
#include "groups.h"
#include "host_threads.h"

#include <algorithm>
#include <climits>
#include <cstring>
#include <thread>

namespace siml {

namespace {

bool row_eligible(const GroupInput& in, int64_t r) {
    const int32_t L = in.len[r];
    if (L > kMaxKeys) return false;
    const int32_t* id = in.ids + in.start[r];
    const int32_t* ct = in.counts + in.start[r];
    int64_t sum = 0;
    for (int32_t k = 0; k < L; ++k) {
        if (ct[k] < 0 || (uint32_t)ct[k] > kCountMax) return false;
        if (k && id[k] <= id[k - 1]) return false;
        sum += ct[k];
    }
    return sum <= (int64_t)kCountMax;
}

// Size of the union of sorted-unique u and v.
int64_t union_size(const std::vector<int32_t>& u, const int32_t* v, int32_t nv) {
    int64_t i = 0, j = 0, n = 0;
    const int64_t nu = (int64_t)u.size();
    while (i < nu && j < nv) {
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
    return n + (nu - i) + (nv - j);
}

void union_into(std::vector<int32_t>& u, const int32_t* v, int32_t nv, std::vector<int32_t>& tmp) {
    tmp.resize(u.size() + (size_t)nv);
    auto e = std::set_union(u.begin(), u.end(), v, v + nv, tmp.begin());
    tmp.resize((size_t)(e - tmp.begin()));
    u.swap(tmp);
}

// xorshift64* -- deterministic multipliers per group
inline uint64_t next_rand(uint64_t& s) {
    s ^= s >> 12;
    s ^= s << 25;
    s ^= s >> 27;
    return s * 2685821657736338717ull;
}

// Places keys[0..K) in a cuckoo table; returns false if the seed fails.
bool place(const std::vector<int32_t>& keys, const GroupHdr& h, std::vector<int32_t>& tab) {
    tab.assign(kSlots, -1);
    const int K = (int)keys.size();
    auto slot = [&](int k, int which) {
        const uint32_t x = (uint32_t)keys[(size_t)k];
        return which ? group_hash(x, h.a2, h.b2) : group_hash(x, h.a1, h.b1);
    };
    for (int k = 0; k < K; ++k) {
        int cur = k;
        uint32_t pos = slot(cur, 0);
        if (pos == 0) pos = slot(cur, 1);
        if (pos == 0) return false;
        bool placed = false;
        for (int it = 0; it < 4 * kSlots; ++it) {
            if (tab[pos] < 0) {
                tab[pos] = cur;
                placed = true;
                break;
            }
            std::swap(cur, tab[pos]);
            const uint32_t p1 = slot(cur, 0), p2 = slot(cur, 1);
            const uint32_t np = (pos == p1) ? p2 : p1;
            if (np == 0 || np == pos) return false;
            pos = np;
        }
        if (!placed) return false;
    }
    return true;
}

struct Plan {
    int64_t row0;
    int32_t nrows;
    int32_t nkeys;
};

}  // namespace

bool build_groups(const GroupInput& in, GroupTables& out, int nthreads) {
    out = GroupTables();
    out.row_group.assign((size_t)in.n, -1);
    // 1. greedy grouping of consecutive eligible rows
    std::vector<Plan> plans;
    std::vector<int32_t> u, tmp;
    Plan cur{0, 0, 0};
    auto flush = [&]() {
        if (cur.nrows > 0) plans.push_back(cur);
        cur = Plan{0, 0, 0};
        u.clear();
    };
    for (int64_t r = 0; r < in.n; ++r) {
        if (!row_eligible(in, r)) {
            flush();
            out.merge_rows.push_back((int32_t)r);
            continue;
        }
        const int32_t* id = in.ids + in.start[r];
        const int32_t L = in.len[r];
        if (cur.nrows == kGroupRows || (cur.nrows > 0 && union_size(u, id, L) > kMaxKeys)) flush();
        if (cur.nrows == 0) cur.row0 = r;
        union_into(u, id, L, tmp);
        cur.nrows += 1;
        cur.nkeys = (int32_t)u.size();
    }
    flush();
    const int64_t G = (int64_t)plans.size();
    out.hdr.resize((size_t)G);
    int64_t total = 0;
    for (int64_t g = 0; g < G; ++g) {
        out.hdr[(size_t)g].entry_off = total;
        total += plans[(size_t)g].nkeys;
    }
    out.ekey.resize((size_t)total);
    out.ecnt.assign((size_t)total * 4, 0u);
    std::vector<char> failed((size_t)std::max<int64_t>(G, 1), 0);

    // 2. per-group union, counts and cuckoo placement (parallel)
    auto work = [&](int64_t g0, int64_t g1) {
        std::vector<int32_t> keys, tab, scratch;
        for (int64_t g = g0; g < g1; ++g) {
            const Plan& p = plans[(size_t)g];
            GroupHdr& h = out.hdr[(size_t)g];
            keys.clear();
            for (int32_t i = 0; i < p.nrows; ++i)
                union_into(keys, in.ids + in.start[p.row0 + i], in.len[p.row0 + i], scratch);
            const int K = (int)keys.size();
            GroupEntry* ek = out.ekey.data() + h.entry_off;
            uint32_t* ec = out.ecnt.data() + h.entry_off * 4;
            for (int32_t i = 0; i < p.nrows; ++i) {
                const int32_t* id = in.ids + in.start[p.row0 + i];
                const int32_t* ct = in.counts + in.start[p.row0 + i];
                int k = 0;
                for (int32_t j = 0; j < in.len[p.row0 + i]; ++j) {
                    while (keys[(size_t)k] < id[j]) ++k;
                    ec[(size_t)k * 4 + (size_t)(i >> 1)] |= (uint32_t)ct[j] << (16 * (i & 1));
                }
            }
            uint64_t rs = 0x9E3779B97F4A7C15ull ^ (uint64_t)(g + 1) * 0xD1B54A32D192ED03ull;
            bool ok = false;
            for (int attempt = 0; attempt < 64 && !ok; ++attempt) {
                h.a1 = (uint32_t)(next_rand(rs) >> 40) | 1u;
                h.b1 = (uint32_t)(next_rand(rs) >> 40) | 1u;
                h.a2 = (uint32_t)(next_rand(rs) >> 40) | 1u;
                h.b2 = (uint32_t)(next_rand(rs) >> 40) | 1u;
                h.pad = 0;
                if (h.a1 == h.a2 && h.b1 == h.b2) continue;
                ok = place(keys, h, tab);
            }
            if (!ok) {
                failed[(size_t)g] = 1;
                continue;
            }
            for (int s = 1; s < kSlots; ++s)
                if (tab[(size_t)s] >= 0) {
                    ek[tab[(size_t)s]].key = keys[(size_t)tab[(size_t)s]];
                    ek[tab[(size_t)s]].slot = s;
                }
            int64_t e = INT32_MIN;
            for (int k = 0; k < K; ++k) {
                if ((int64_t)keys[(size_t)k] == e)
                    ++e;
                else if ((int64_t)keys[(size_t)k] > e)
                    break;
            }
            h.row0 = (int32_t)p.row0;
            h.nrows = p.nrows;
            h.nkeys = K;
            h.empty_key = (int32_t)e;
        }
    };
    const int nt = nthreads > 0 ? nthreads : default_threads(32);
    if (G < 256 || nt <= 1) {
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

    // 3. groups whose table could not be built go to the merge path
    std::vector<GroupHdr> hdr;
    std::vector<GroupEntry> ekey;
    std::vector<uint32_t> ecnt;
    bool any_failed = false;
    for (int64_t g = 0; g < G; ++g) any_failed |= failed[(size_t)g] != 0;
    if (any_failed) {
        for (int64_t g = 0; g < G; ++g) {
            const GroupHdr& h = out.hdr[(size_t)g];
            const Plan& p = plans[(size_t)g];
            if (failed[(size_t)g]) {
                for (int32_t i = 0; i < p.nrows; ++i) out.merge_rows.push_back((int32_t)(p.row0 + i));
                continue;
            }
            GroupHdr nh = h;
            nh.entry_off = (int64_t)ekey.size();
            ekey.insert(ekey.end(), out.ekey.begin() + h.entry_off, out.ekey.begin() + h.entry_off + h.nkeys);
            ecnt.insert(ecnt.end(), out.ecnt.begin() + h.entry_off * 4, out.ecnt.begin() + (h.entry_off + h.nkeys) * 4);
            hdr.push_back(nh);
        }
        out.hdr.swap(hdr);
        out.ekey.swap(ekey);
        out.ecnt.swap(ecnt);
        std::sort(out.merge_rows.begin(), out.merge_rows.end());
    }
    for (size_t g = 0; g < out.hdr.size(); ++g)
        for (int32_t i = 0; i < out.hdr[g].nrows; ++i) out.row_group[(size_t)(out.hdr[g].row0 + i)] = (int32_t)g;
    return true;
}

}  // namespace siml
