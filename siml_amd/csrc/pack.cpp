// pack.cpp -- host-side SMILES -> SIML set packer (part of libsiml_amd.so).
//
// Behaviour follows _ccompiler.SMILEStoMatrices (_ccompiler.cpp:276-381):
//   1. digit normalisation (processSMILESdigits, _ccompiler.cpp:178-264),
//      written here as a character-class x state transition table;
//   2. every 4-byte window of the normalised string as a little-endian int32
//      (LingoBag, _ccompiler.cpp:83-165), magnitude = length-3 (0 below 4);
//   3. sorted (signed int32) unique ids with multiplicities.  The reference
//      inserts one window at a time into a sorted array (:48-77); here the
//      windows are sorted once and run-length encoded, which gives the same
//      multiset.
// The reference reads one stale byte past its zeroed buffer when a string
// of length 0 mod 4 has no terminator (:120-125, :304); only real windows
// are produced here.
#include "host_threads.h"

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "siml_internal.h"

namespace {

enum CharClass : uint8_t { C_OTHER = 0, C_DIGIT, C_PERCENT, C_SPECIAL, C_STOP };
enum State : uint8_t { S_NORMAL = 0, S_REPLACE, S_CHOMP, S_SPECIAL, S_DONE };

struct ClassTable {
    uint8_t cls[256];
    ClassTable() {
        std::memset(cls, C_OTHER, sizeof(cls));
        for (int c = '0'; c <= '9'; ++c) cls[c] = C_DIGIT;
        cls[(unsigned char)'%'] = C_PERCENT;
        cls[(unsigned char)'+'] = cls[(unsigned char)'-'] = cls[(unsigned char)'H'] = cls[(unsigned char)'['] =
            C_SPECIAL;
        cls[(unsigned char)' '] = cls[(unsigned char)'\n'] = cls[(unsigned char)'\r'] = C_STOP;
    }
};
const ClassTable kClass;

// Action per (state, class): what to emit and where to go.
//   emit: 0 = nothing, 1 = the input byte, 2 = '0'
struct Step {
    uint8_t emit, next;
};
// Rows: state; columns: class (OTHER, DIGIT, PERCENT, SPECIAL, STOP).
constexpr Step kTable[4][5] = {
    // NORMAL (:229-256)
    {{1, S_NORMAL}, {2, S_REPLACE}, {1, S_CHOMP}, {1, S_SPECIAL}, {0, S_DONE}},
    // REPLACE (:198-221): + - H [ are plain copies here
    {{1, S_NORMAL}, {2, S_REPLACE}, {1, S_CHOMP}, {1, S_NORMAL}, {0, S_DONE}},
    // CHOMP (:194-197): the byte is dropped whatever it is
    {{0, S_REPLACE}, {0, S_REPLACE}, {0, S_REPLACE}, {0, S_REPLACE}, {0, S_REPLACE}},
    // SPECIALNUMS (:222-228): verbatim copy, stay while digits
    {{1, S_NORMAL}, {1, S_SPECIAL}, {1, S_NORMAL}, {1, S_NORMAL}, {1, S_NORMAL}},
};

// Normalise one string into `out` (capacity >= len); returns its length.
inline int64_t normalise(const char* s, int64_t len, char* out) {
    uint8_t st = S_NORMAL;
    int64_t n = 0;
    for (int64_t p = 0; p < len; ++p) {
        const unsigned char c = (unsigned char)s[p];
        if (c == 0) break;
        const Step a = kTable[st][kClass.cls[c]];
        if (a.emit == 1)
            out[n++] = (char)c;
        else if (a.emit == 2)
            out[n++] = '0';
        st = a.next;
        if (st == S_DONE) break;
    }
    return n;
}

struct Bag {
    std::vector<char> norm;
    std::vector<int32_t> win;
    std::vector<int32_t> ids, cnt;
    int32_t mag = 0;

    void build(const char* s, int64_t len) {
        norm.resize((size_t)len + 4);
        const int64_t n = normalise(s, len, norm.data());
        ids.clear();
        cnt.clear();
        if (n < 4) {
            mag = 0;
            return;
        }
        win.resize((size_t)(n - 3));
        const unsigned char* u = (const unsigned char*)norm.data();
        for (int64_t k = 0; k + 3 < n; ++k)
            win[(size_t)k] = (int32_t)((uint32_t)u[k] | ((uint32_t)u[k + 1] << 8) | ((uint32_t)u[k + 2] << 16) |
                                       ((uint32_t)u[k + 3] << 24));
        std::sort(win.begin(), win.end());
        for (size_t k = 0; k < win.size();) {
            size_t e = k + 1;
            while (e < win.size() && win[e] == win[k]) ++e;
            ids.push_back(win[k]);
            cnt.push_back((int32_t)(e - k));
            k = e;
        }
        mag = (int32_t)(n - 3);
    }
};

template <class F>
void parallel_rows(int64_t n, int nthreads, F&& f) {
    const int nt = nthreads > 0 ? std::min(nthreads, 64) : siml::default_threads(64);
    if (n < 4096 || nt <= 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    const int64_t chunk = (n + nt - 1) / nt;
    for (int t = 0; t < nt; ++t) {
        const int64_t a = t * chunk, b = std::min(n, a + chunk);
        if (a >= b) break;
        th.emplace_back([&f, a, b] { f(a, b); });
    }
    for (auto& t : th) t.join();
}

}  // namespace

extern "C" int siml_pack_measure(const char* buf, const int64_t* offsets, int64_t n, int32_t* lengths,
                                 int32_t* mags, int nthreads) {
    if (n < 0 || (n > 0 && (!buf || !offsets || !lengths || !mags)))
        return siml::fail(SIML_EINVAL, "siml_pack_measure: null argument or negative count");
    for (int64_t r = 0; r < n; ++r)
        if (offsets[r + 1] < offsets[r]) return siml::fail(SIML_EINVAL, "siml_pack_measure: offsets not ascending");
    parallel_rows(n, nthreads, [&](int64_t a, int64_t b) {
        Bag bag;
        for (int64_t r = a; r < b; ++r) {
            bag.build(buf + offsets[r], offsets[r + 1] - offsets[r]);
            lengths[r] = (int32_t)bag.ids.size();
            mags[r] = bag.mag;
        }
    });
    return SIML_OK;
}

extern "C" int siml_pack_fill(const char* buf, const int64_t* offsets, int64_t n, int32_t* lingos, int32_t* counts,
                              int64_t width, int nthreads) {
    if (n < 0 || width < 0 || (n > 0 && (!buf || !offsets || ((!lingos || !counts) && width > 0))))
        return siml::fail(SIML_EINVAL, "siml_pack_fill: null argument or negative size");
    std::vector<int> bad((size_t)std::max<int64_t>(n, 1), 0);
    parallel_rows(n, nthreads, [&](int64_t a, int64_t b) {
        Bag bag;
        for (int64_t r = a; r < b; ++r) {
            bag.build(buf + offsets[r], offsets[r + 1] - offsets[r]);
            const int64_t L = (int64_t)bag.ids.size();
            if (L > width) {
                bad[(size_t)r] = 1;
                continue;
            }
            if (width == 0) continue;
            int32_t* lr = lingos + r * width;
            int32_t* cr = counts + r * width;
            std::memset(lr, 0, sizeof(int32_t) * (size_t)width);
            std::memset(cr, 0, sizeof(int32_t) * (size_t)width);
            if (L) {
                std::memcpy(lr, bag.ids.data(), sizeof(int32_t) * (size_t)L);
                std::memcpy(cr, bag.cnt.data(), sizeof(int32_t) * (size_t)L);
            }
        }
    });
    for (int64_t r = 0; r < n; ++r)
        if (bad[(size_t)r]) return siml::fail(SIML_EINVAL, "siml_pack_fill: a row has more Lingos than width");
    return SIML_OK;
}

extern "C" int siml_pack_fill_ragged(const char* buf, const int64_t* offsets, int64_t n, const int64_t* row_offsets,
                                     int32_t* ids, int32_t* counts, int nthreads) {
    if (n < 0 || (n > 0 && (!buf || !offsets || !row_offsets)))
        return siml::fail(SIML_EINVAL, "siml_pack_fill_ragged: null argument or negative count");
    std::vector<int> bad((size_t)std::max<int64_t>(n, 1), 0);
    parallel_rows(n, nthreads, [&](int64_t a, int64_t b) {
        Bag bag;
        for (int64_t r = a; r < b; ++r) {
            bag.build(buf + offsets[r], offsets[r + 1] - offsets[r]);
            const int64_t L = (int64_t)bag.ids.size();
            if (row_offsets[r + 1] - row_offsets[r] != L) {
                bad[(size_t)r] = 1;
                continue;
            }
            if (L) {
                std::memcpy(ids + row_offsets[r], bag.ids.data(), sizeof(int32_t) * (size_t)L);
                std::memcpy(counts + row_offsets[r], bag.cnt.data(), sizeof(int32_t) * (size_t)L);
            }
        }
    });
    for (int64_t r = 0; r < n; ++r)
        if (bad[(size_t)r]) return siml::fail(SIML_EINVAL, "siml_pack_fill_ragged: row_offsets do not match lengths");
    return SIML_OK;
}
