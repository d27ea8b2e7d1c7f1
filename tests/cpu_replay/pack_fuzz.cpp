// Randomised run of the SMILES packer's C entry points (test
// infrastructure, built with AddressSanitizer and UndefinedBehaviorSanitizer
// by tests/test_sanitizers_cpu.py).
//
// Random byte strings (SMILES characters, ring-closure digits, '%' forms and
// arbitrary bytes including NUL and bytes above 127) go through
// siml_pack_measure, siml_pack_fill and siml_pack_fill_ragged with one and
// with several threads.  Checked per string: ids strictly ascending, counts
// positive, counts summing to the magnitude, padded rows equal to ragged
// rows with zero padding, results independent of the thread count, and the
// documented errors for a too-narrow width or mismatched row offsets.
//
// usage: pack_fuzz SEED NSTRINGS
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "siml_internal.h"

namespace siml {
// The library defines this next to its device code; the packer alone needs
// only the return code.
int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    va_end(ap);
    return code;
}
}  // namespace siml

static std::string random_string(std::mt19937& rng) {
    static const char kAlpha[] = "CNOSPFIcnosBrl()=#[]@+-/\\.123456789%0H";
    const int len = (int)(rng() % 160);
    std::string s;
    for (int i = 0; i < len; ++i) {
        const unsigned r = rng() % 100;
        if (r < 80)
            s.push_back(kAlpha[rng() % (sizeof(kAlpha) - 1)]);
        else if (r < 95)
            s.push_back((char)('0' + rng() % 10));
        else
            s.push_back((char)(rng() % 256));
    }
    return s;
}

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: pack_fuzz SEED NSTRINGS\n");
        return 2;
    }
    std::mt19937 rng((unsigned)std::atoi(argv[1]));
    const int64_t n = std::atoll(argv[2]);
    std::string buf;
    std::vector<int64_t> offs{0};
    for (int64_t i = 0; i < n; ++i) {
        buf += random_string(rng);
        offs.push_back((int64_t)buf.size());
    }
    // exact-size heap copy, so a read past the last string is caught
    std::vector<char> bytes(buf.begin(), buf.end());
    const char* b = bytes.empty() ? "" : bytes.data();
    int64_t differ = 0;
    auto check = [&](bool ok, const char* what, int64_t r) {
        if (!ok) {
            if (differ < 10) std::printf("string %lld: %s\n", (long long)r, what);
            ++differ;
        }
    };

    std::vector<int32_t> len1(n), mag1(n), len4(n), mag4(n);
    check(siml_pack_measure(b, offs.data(), n, len1.data(), mag1.data(), 1) == SIML_OK, "measure failed", -1);
    check(siml_pack_measure(b, offs.data(), n, len4.data(), mag4.data(), 4) == SIML_OK, "measure failed", -1);
    check(len1 == len4 && mag1 == mag4, "measure depends on threads", -1);
    std::vector<int64_t> roff(n + 1, 0);
    int32_t width = 0;
    for (int64_t r = 0; r < n; ++r) {
        roff[r + 1] = roff[r] + len1[r];
        width = std::max(width, len1[r]);
    }
    const int64_t nnz = roff[n];
    std::vector<int32_t> ids(nnz + 1), cnt(nnz + 1);
    check(siml_pack_fill_ragged(b, offs.data(), n, roff.data(), ids.data(), cnt.data(), 3) == SIML_OK,
          "ragged fill failed", -1);
    std::vector<int32_t> lp((size_t)(n * width + 1)), cp((size_t)(n * width + 1));
    check(siml_pack_fill(b, offs.data(), n, lp.data(), cp.data(), width, 2) == SIML_OK, "padded fill failed", -1);
    for (int64_t r = 0; r < n; ++r) {
        int64_t sum = 0;
        for (int64_t j = roff[r]; j < roff[r + 1]; ++j) {
            check(cnt[j] > 0, "count not positive", r);
            if (j > roff[r]) check(ids[j - 1] < ids[j], "ids not strictly ascending", r);
            sum += cnt[j];
        }
        check(sum == mag1[r], "counts do not sum to the magnitude", r);
        for (int32_t c = 0; c < width; ++c) {
            const int64_t j = roff[r] + c;
            const bool in = c < len1[r];
            check(lp[r * width + c] == (in ? ids[j] : 0) && cp[r * width + c] == (in ? cnt[j] : 0),
                  "padded row differs from ragged row", r);
        }
    }
    if (width > 0) {
        std::vector<int32_t> lq((size_t)(n * (width - 1) + 1)), cq((size_t)(n * (width - 1) + 1));
        check(siml_pack_fill(b, offs.data(), n, lq.data(), cq.data(), width - 1, 2) == SIML_EINVAL,
              "narrow width not rejected", -1);
    }
    if (n > 0 && len1[0] > 0) {
        std::vector<int64_t> bad = roff;
        for (int64_t r = 1; r <= n; ++r) bad[r] -= 1;
        check(siml_pack_fill_ragged(b, offs.data(), n, bad.data(), ids.data(), cnt.data(), 2) == SIML_EINVAL,
              "mismatched row offsets not rejected", -1);
    }
    check(siml_pack_measure(nullptr, offs.data(), 1, len1.data(), mag1.data(), 1) == SIML_EINVAL,
          "null buffer not rejected", -1);
    std::printf("%lld strings, %lld bytes, %lld ids, %lld differ\n", (long long)n, (long long)buf.size(),
                (long long)nnz, (long long)differ);
    return differ == 0 ? 0 : 1;
}
