// The <ply>.kdtree.radii cache: the text serialization archive of one std::vector<float>
// (main.cu:115-127). Layout written, token by token (space separated, newline at the end):
//   22 serialization::archive   the length-prefixed archive signature
//   17                          archive library version (Boost 1.71)
//   0 0                         class tracking level and class version of std::vector<float>
//   <n>                         element count
//   0                           item version
//   v0 ... v(n-1)               each value as printf "%.9e" (9 fractional digits, enough to
//                               round-trip every float exactly)
// The reader accepts any library version: versions above 3 carry the item version, older ones
// do not (the rule of Boost's collection loader); the count must equal the values that follow.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string_view>

#include "rm_io.h"

namespace rmh {

namespace {

bool is_int_token(std::string_view t) {
    if (t.empty()) return false;
    size_t k = (t[0] == '-' || t[0] == '+') ? 1 : 0;
    if (k == t.size()) return false;
    for (; k < t.size(); ++k)
        if (t[k] < '0' || t[k] > '9') return false;
    return true;
}

}  // namespace

void write_radii(const std::string& path, const float* radii, int64_t n) {
    if (n < 0) throw std::runtime_error("write_radii: negative count");
    std::FILE* f = std::fopen(path.c_str(), "w");
    if (!f) throw std::runtime_error("could not open '" + path + "' for writing");
    std::fprintf(f, "22 serialization::archive 17 0 0 %lld 0", (long long)n);
    for (int64_t i = 0; i < n; ++i) std::fprintf(f, " %.9e", (double)radii[i]);
    std::fputc('\n', f);
    const bool bad = std::ferror(f) != 0;
    if (std::fclose(f) != 0 || bad) throw std::runtime_error("write failed for '" + path + "'");
}

std::vector<float> read_radii(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("could not open radii file '" + path + "'");
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    // whitespace-separated tokens, scanned in place (a 10M-point cache is ~160 MB of text)
    size_t pos = 0;
    auto next = [&](std::string_view& tok) {
        while (pos < text.size() && std::isspace((unsigned char)text[pos])) ++pos;
        const size_t b = pos;
        while (pos < text.size() && !std::isspace((unsigned char)text[pos])) ++pos;
        tok = std::string_view(text.data() + b, pos - b);
        return !tok.empty();
    };
    const std::string bad = "radii file '" + path + "' is not a serialization archive of floats";
    std::string_view t0, t1, t2;
    if (!next(t0) || !next(t1) || !next(t2) || t0 != "22" || t1 != "serialization::archive" ||
        !is_int_token(t2))
        throw std::runtime_error(bad + " (invalid signature)");
    // Header after the signature and library version: class tracking level, class version,
    // element count, then (library versions above 3) the item version.
    const long long lib_version = std::atoll(std::string(t2).c_str());
    const int h = lib_version > 3 ? 4 : 3;
    long long count = -1;
    for (int k = 0; k < h; ++k) {
        std::string_view t;
        if (!next(t)) throw std::runtime_error(bad + " (truncated header)");
        if (!is_int_token(t)) throw std::runtime_error(bad + " (bad header token '" + std::string(t) + "')");
        if (k == 2) count = std::atoll(std::string(t).c_str());
    }
    if (count < 0) throw std::runtime_error(bad + " (negative count)");
    std::vector<float> out;
    out.reserve((size_t)std::min<long long>(count, (long long)(text.size() / 2 + 1)));
    std::string_view t;
    while (next(t)) {
        const std::string tok(t);
        char* end = nullptr;
        const float v = std::strtof(tok.c_str(), &end);
        if (end != tok.c_str() + tok.size()) throw std::runtime_error(bad + " (bad value '" + tok + "')");
        out.push_back(v);
    }
    if ((long long)out.size() != count)
        throw std::runtime_error(bad + " (holds " + std::to_string(out.size()) +
                                 " values, its header says " + std::to_string(count) + ")");
    return out;
}

}  // namespace rmh
