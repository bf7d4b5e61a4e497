// Text formats of the reference's command-line driver: whitespace-separated
// float matrices in (src/Utilities.cpp:41-60), one line of ids per query out
// (src/Utilities.cpp:13-31), and the file name a verbose query writes
// (src/FalconnPP.cpp:694-700).  Host code only.
#include "textio.h"

#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "common.h"

namespace fpp {

namespace {

bool is_space(char c) { return std::isspace(static_cast<unsigned char>(c)) != 0; }

// Converts the token [b, e) (e is writable and holds a space) as fscanf's %f
// does.  Returns false when no number starts the token; *whole tells whether
// the number took the whole token (if not, a later read would fail there).
bool parse_token(char* b, char* e, float& v, bool& whole) {
    const char keep = *e;
    *e = '\0';
    char* end = nullptr;
    errno = 0;
    v = std::strtof(b, &end);
    *e = keep;
    if (end == b) return false;
    whole = end == e;
    return true;
}

}  // namespace

int64_t load_txt(const char* path, int64_t count, float* out) {
    FILE* f = std::fopen(path, "rb");
    FPP_REQUIRE(f != nullptr, FPP_EIO, std::string("cannot open data file ") + path);
    for (int64_t i = 0; i < count; ++i) out[i] = 0.0f;
    constexpr size_t CHUNK = size_t(1) << 24;
    std::vector<char> buf(CHUNK + 2);
    size_t carry = 0;  // bytes of an unfinished token kept at the front
    int64_t n = 0;
    bool stop = false, eof = false;
    while (!stop && n < count && !eof) {
        const size_t got = std::fread(buf.data() + carry, 1, CHUNK - carry, f);
        eof = got < CHUNK - carry;
        const size_t len = carry + got;
        buf[len] = ' ';
        // the last token may continue in the next chunk unless this is the end
        size_t limit = len;
        if (!eof) {
            while (limit > 0 && !is_space(buf[limit - 1])) --limit;
            if (limit == 0) {
                std::fclose(f);
                throw Error(FPP_EIO, std::string("token longer than 16 MiB in ") + path);
            }
        }
        size_t i = 0;
        while (i < limit && n < count) {
            while (i < limit && is_space(buf[i])) ++i;
            if (i >= limit) break;
            size_t j = i;
            while (j < limit && !is_space(buf[j])) ++j;
            float v;
            bool whole;
            if (!parse_token(buf.data() + i, buf.data() + j, v, whole)) {
                stop = true;
                break;
            }
            out[n++] = v;
            if (!whole) {
                stop = true;
                break;
            }
            i = j;
        }
        carry = len - limit;
        if (carry) std::memmove(buf.data(), buf.data() + limit, carry);
    }
    std::fclose(f);
    return n;
}

void write_topk_txt(const char* path, const int32_t* ids, int64_t nq, int k) {
    FILE* f = std::fopen(path, "w");
    FPP_REQUIRE(f != nullptr, FPP_EIO, std::string("cannot create output file ") + path);
    bool ok = true;
    for (int64_t q = 0; q < nq && ok; ++q) {
        for (int i = 0; i < k && ok; ++i) ok = std::fprintf(f, "%d ", (int)ids[q * k + i]) > 0;
        ok = ok && std::fputc('\n', f) != EOF;
    }
    ok = (std::fclose(f) == 0) && ok;
    FPP_REQUIRE(ok, FPP_EIO, std::string("cannot write output file ") + path);
}

std::string topk_file_name(int k, int n_proj, int n_tables, int iProbes, int bucket_minSize, float bucket_scale,
                           int qProbes, const char* layout) {
    // the scale enters as (int)(scale * 100), a float product truncated
    const float pct = bucket_scale * 100;
    return std::string(layout) + "_Top_" + std::to_string(k) + "_D_" + std::to_string(n_proj) + "_L_" + std::to_string(n_tables) +
           "_iProbe_" + std::to_string(iProbes) + "_bucketMinSize_" + std::to_string(bucket_minSize) +
           "_bucketScale_" + std::to_string((int)pct) + "_qProbe_" + std::to_string(qProbes) + ".txt";
}

}  // namespace fpp
