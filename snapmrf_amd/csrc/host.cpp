// host.cpp -- host-side boundary of the snapMRF path: error state, rfpulse CSV
// reader, T1/T2/B0/B1 grid parser and RA file I/O.  Pure C++17, no HIP.
//
// Float32 arithmetic is kept exactly as in the reference host code (this file
// is compiled with -ffp-contract=off so no multiply-add is fused), so the grid
// values and atom order are bit-identical to src/functions.cu:196-414.
#include "snapmrf.h"
#include "internal.h"

#include <cerrno>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

namespace {
thread_local char g_err[1024] = "";
}

extern "C" const char *mrf_last_error(void) { return g_err; }

int mrf_fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

void mrf_clear_error(void) { g_err[0] = '\0'; }

// ------------------------------------------------------------------ CSV ----

extern "C" size_t mrf_csv_dim(const char *filepath)
{
    // src/functions.cu:100-120: count '\n'; the loop variable is a (signed)
    // char compared with EOF, so a 0xFF byte also ends the count.
    FILE *f = filepath ? std::fopen(filepath, "r") : nullptr;
    if (!f) {
        mrf_fail(MRF_E_IO, "unable to open file %s", filepath ? filepath : "(null)");
        return (size_t)-1;
    }
    size_t rows = 0;
    for (char c = (char)std::getc(f); c != (char)EOF; c = (char)std::getc(f))
        if (c == '\n') rows++;
    std::fclose(f);
    return rows;
}

extern "C" int mrf_csv_read(float *h_mrf, size_t nreps, const char *filepath)
{
    mrf_clear_error();
    if (!h_mrf && nreps) return mrf_fail(MRF_E_ARG, "mrf_csv_read: null output");
    FILE *f = filepath ? std::fopen(filepath, "r") : nullptr;
    if (!f) return mrf_fail(MRF_E_IO, "unable to open file %s", filepath ? filepath : "(null)");
    // src/functions.cu:122-158: lines are consumed as fgets(buf, 100) chunks
    // and the first two chunks are skipped (a 99-character header plus its
    // newline).  A shorter header makes the second chunk a data row, which
    // the reference silently drops; we do the same and say so.
    char buf[100];
    size_t index = 0, count = 0;
    bool header_short = false;
    while (index < nreps && std::fgets(buf, sizeof(buf), f)) {
        if (count == 0 && std::strchr(buf, '\n')) header_short = true;
        if (count > 1) {
            char *save = nullptr;
            char *tok = strtok_r(buf, ",", &save);
            for (int c = 0; c < 4; c++) {
                if (!tok) {
                    std::fclose(f);
                    return mrf_fail(MRF_E_FORMAT, "%s: row %zu has fewer than 4 columns",
                                    filepath, index + 1);
                }
                h_mrf[(size_t)c * nreps + index] = (float)std::atof(tok);
                tok = strtok_r(nullptr, ",", &save);
            }
            index++;
        }
        count++;
    }
    std::fclose(f);
    if (index < nreps) {
        for (size_t i = index; i < nreps; i++)
            for (int c = 0; c < 4; c++) h_mrf[(size_t)c * nreps + i] = 0.0f;
        mrf_fail(MRF_OK, "%s: %zu of %zu rows read%s; the rest were zero-filled",
                 filepath, index, nreps,
                 header_short ? " (header shorter than 99 characters: its next line was "
                                "consumed as the second skipped chunk)"
                              : "");
        return 1;
    }
    return MRF_OK;
}

// ----------------------------------------------------------------- grids ---

namespace {

// parse_range, src/functions.cu:196-235.  atof() results are stored in floats.
bool parse_range(const char *str, float *mn, float *step, float *mx, size_t *len)
{
    float tmp = (float)std::atof(str);
    *mn = tmp;
    const char *p = std::strchr(str, ':');
    if (!p) {
        *step = 0.0f;
        *mx = tmp;
        *len = 1;
        return true;
    }
    p = p + 1;
    tmp = (float)std::atof(p);
    *step = tmp;
    p = std::strchr(p, ':');
    if (!p) {
        *step = 1.0f;
        *mx = tmp;
        if (!(*mx >= *mn)) return false;   // reference: assert
        *len = (size_t)(((*mx - *mn) / 1) + 1);
    } else {
        p = p + 1;
        tmp = (float)std::atof(p);
        *mx = tmp;
        if (!(*mx >= *mn)) return false;   // reference: assert
        float q = ((*mx - *mn) / *step) + 1;
        if (!(q >= 1.0f) || !std::isfinite(q)) return false;
        *len = (size_t)q;
    }
    return true;
}

// Shared walk of parse_length / parse_params (src/functions.cu:237-353).
// The comparison |tmp - min| < 1e-7 uses the float overload of abs, and the
// running end `tmp` is min + len*step after the first range (one step past the
// end when the range was not deduplicated) -- both kept as in the reference.
bool grid_walk(const char *str, float *out, size_t *total)
{
    if (!str) return false;
    size_t n = std::strlen(str);
    if (n >= 1024) return false;   // reference copies into char[1024]
    char cpy[1024];
    std::memcpy(cpy, str, n + 1);
    float mn = 0, step = 0, mx = 0, tmp = 0;
    size_t len = 0, stride = 0;
    int count = 0;
    char *save = nullptr;
    for (char *r = strtok_r(cpy, "+", &save); r; r = strtok_r(nullptr, "+", &save)) {
        count++;
        if (!parse_range(r, &mn, &step, &mx, &len)) return false;
        if (count == 1) {
            tmp = mn + (float)(len - 1) * step;
            if (out)
                for (size_t i = 0; i < len; i++) out[stride + i] = mn + (float)(int)i * step;
        } else {
            if (std::fabs(tmp - mn) < 1E-7) {
                len = len - 1;
                if (out)
                    for (size_t i = 0; i < len; i++)
                        out[stride + i] = (mn + step) + (float)(int)i * step;
            } else if (out) {
                for (size_t i = 0; i < len; i++) out[stride + i] = mn + (float)(int)i * step;
            }
            tmp = mn + (float)len * step;
        }
        stride += len;
    }
    if (count == 0) return false;
    *total = stride;
    return true;
}

}  // namespace

extern "C" size_t mrf_parse_length(const char *str)
{
    size_t total = 0;
    if (!grid_walk(str, nullptr, &total)) {
        mrf_fail(MRF_E_ARG, "malformed grid string '%s'", str ? str : "(null)");
        return 0;
    }
    return total;
}

extern "C" int mrf_parse_params(float *h_p, const char *str, size_t len)
{
    size_t total = 0;
    if (!h_p) return mrf_fail(MRF_E_ARG, "mrf_parse_params: null output");
    if (!grid_walk(str, nullptr, &total))
        return mrf_fail(MRF_E_ARG, "malformed grid string '%s'", str ? str : "(null)");
    if (total != len)
        return mrf_fail(MRF_E_ARG, "grid '%s' has %zu values, %zu given", str, total, len);
    grid_walk(str, h_p, &total);
    return MRF_OK;
}

extern "C" void mrf_logspace(float *h_p, float mn, float mx, size_t len)
{
    // src/functions.cu:269-304.  pow/log10 resolve to the float overloads in
    // the reference's C++ translation unit; the symmetric branch writes
    // h_p[0 .. 2*(len/2)+1], one past len for odd len.
    if (mn > mx) { float t = mn; mn = mx; mx = t; }
    if (len == 0) return;
    if (len == 1) {
        h_p[0] = (mn + mx) / 2.f;
        return;
    }
    if (mn > 0.f) {
        float logmin = std::log10(mn), logmax = std::log10(mx);
        float logstep = (logmax - logmin) / (float)(len - 1);
        for (size_t i = 0; i < len; i++)
            h_p[i] = std::pow(10.f, logmin + (float)(int)i * logstep);
    } else {
        float logmax = std::log10(mx);
        size_t h = len / 2;
        float logstep = logmax / (float)h;
        for (size_t i = 0; i < h + 1; i++) h_p[i] = -std::pow(10.f, (float)(h - i) * logstep);
        h_p[h] = 0.f;
        for (size_t i = 0; i < h + 1; i++) h_p[i + h + 1] = std::pow(10.f, (float)(i + 1) * logstep);
    }
}

extern "C" size_t mrf_compute_natoms(const float *t1, const float *t2, size_t l_t1,
                                     size_t l_t2, size_t l_b0, size_t l_b1)
{
    size_t n = 0;   // src/functions.cu:393-414
    for (size_t i = 0; i < l_t1; i++)
        for (size_t j = 0; j < l_t2; j++)
            if (t1[i] > t2[j]) n++;
    return n * l_b0 * l_b1;
}

extern "C" int mrf_trans_params(float *p, const float *t1, const float *t2, const float *b0,
                                const float *b1, size_t l_t1, size_t l_t2, size_t l_b0,
                                size_t l_b1, size_t NATOMS, int nparams)
{
    if (nparams < 4) return mrf_fail(MRF_E_ARG, "nparams must be >= 4");
    size_t count = 0;   // src/functions.cu:355-391: B1 outer, then B0, T2, T1 fastest
    for (size_t h = 0; h < l_b1; h++)
        for (size_t i = 0; i < l_b0; i++)
            for (size_t j = 0; j < l_t2; j++)
                for (size_t k = 0; k < l_t1; k++)
                    if (t1[k] > t2[j]) {
                        if (count >= NATOMS)
                            return mrf_fail(MRF_E_ARG, "NATOMS %zu too small", NATOMS);
                        float *q = p + count * (size_t)nparams;
                        q[0] = t1[k];
                        q[1] = t2[j];
                        q[2] = b0[i];
                        q[3] = b1[h];
                        count++;
                    }
    if (count != NATOMS)
        return mrf_fail(MRF_E_ARG, "grid has %zu atoms, NATOMS = %zu", count, NATOMS);
    return MRF_OK;
}

extern "C" int mrf_grid_natoms_mode(size_t N, size_t *l_t1, size_t *l_t2, size_t *l_b0,
                                    size_t *l_b1, float *t1, float *t2, float *b0, float *b1)
{
    // src/mrf.cu:179-203
    *l_b1 = 11;
    *l_b0 = 21;
    size_t lt = (size_t)std::roundf(std::sqrt((float)(N / *l_b0 / *l_b1)));
    if (lt == 0) lt = 1;
    *l_t1 = lt;
    *l_t2 = lt;
    if (t1) mrf_logspace(t1, 100.f, 3000.f, lt);
    if (t2) mrf_logspace(t2, 20.f, 1000.f, lt);
    if (b0) mrf_logspace(b0, -150.f, 150.f, *l_b0);
    if (b1) mrf_logspace(b1, 0.5f, 1.5f, *l_b1);
    return MRF_OK;
}

// ------------------------------------------------------------------- RA ----

namespace {
const uint64_t kRaMagic = 0x7961727261776172ULL;
const uint64_t kRaChunk = 1ULL << 30;

bool read_full(int fd, void *dst, uint64_t n)
{
    uint8_t *p = (uint8_t *)dst;
    while (n) {
        ssize_t r = ::read(fd, p, (size_t)(n < kRaChunk ? n : kRaChunk));
        if (r < 0 && errno == EINTR) continue;
        if (r <= 0) return false;
        p += r;
        n -= (uint64_t)r;
    }
    return true;
}

bool write_full(int fd, const void *src, uint64_t n)
{
    const uint8_t *p = (const uint8_t *)src;
    while (n) {
        ssize_t r = ::write(fd, p, (size_t)(n < kRaChunk ? n : kRaChunk));
        if (r < 0 && errno == EINTR) continue;
        if (r <= 0) return false;
        p += r;
        n -= (uint64_t)r;
    }
    return true;
}
}  // namespace

extern "C" int mrf_ra_read(mrf_ra_t *a, const char *path)
{
    // src/ra.cu:23-55, with the header validated and short reads reported.
    if (!a || !path) return mrf_fail(MRF_E_ARG, "mrf_ra_read: null argument");
    std::memset(a, 0, sizeof(*a));
    int fd = ::open(path, O_RDONLY);
    if (fd < 0) return mrf_fail(MRF_E_IO, "unable to open %s for reading", path);
    uint64_t h[6];
    if (!read_full(fd, h, sizeof(h))) {
        ::close(fd);
        return mrf_fail(MRF_E_FORMAT, "%s: truncated RA header", path);
    }
    if (h[0] != kRaMagic) {
        ::close(fd);
        return mrf_fail(MRF_E_FORMAT, "%s: not an RA file (bad magic)", path);
    }
    if (h[1] & 1ULL) {
        ::close(fd);
        return mrf_fail(MRF_E_UNSUPPORTED, "%s: big-endian RA files are not supported", path);
    }
    a->flags = h[1];
    a->eltype = h[2];
    a->elbyte = h[3];
    a->size = h[4];
    a->ndims = h[5];
    if (a->ndims == 0 || a->ndims > 64) {
        ::close(fd);
        return mrf_fail(MRF_E_FORMAT, "%s: bad ndims %llu", path, (unsigned long long)a->ndims);
    }
    a->dims = (uint64_t *)std::malloc(a->ndims * sizeof(uint64_t));
    if (!a->dims || !read_full(fd, a->dims, a->ndims * sizeof(uint64_t))) {
        ::close(fd);
        mrf_ra_free(a);
        return mrf_fail(MRF_E_FORMAT, "%s: truncated RA dims", path);
    }
    // size must be elbyte * prod(dims) (src/ra.h:20-58), computed without
    // overflow: callers size their copies from the dims
    uint64_t want = a->elbyte;
    bool overflow = false;
    for (uint64_t d = 0; d < a->ndims; d++) {
        if (a->dims[d] && want > UINT64_MAX / a->dims[d]) overflow = true;
        want *= a->dims[d];
    }
    if (overflow || want != a->size) {
        ::close(fd);
        const unsigned long long sz = (unsigned long long)a->size;
        mrf_ra_free(a);
        return mrf_fail(MRF_E_FORMAT, "%s: header size %llu is not elbyte x prod(dims)%s", path,
                        sz, overflow ? " (dims overflow)" : "");
    }
    a->data = (uint8_t *)std::malloc(a->size ? a->size : 1);
    if (!a->data) {
        ::close(fd);
        mrf_ra_free(a);
        return mrf_fail(MRF_E_NOMEM, "%s: cannot allocate %llu bytes", path,
                        (unsigned long long)a->size);
    }
    if (!read_full(fd, a->data, a->size)) {
        ::close(fd);
        mrf_ra_free(a);
        return mrf_fail(MRF_E_FORMAT, "%s: data shorter than header size", path);
    }
    ::close(fd);
    return MRF_OK;
}

extern "C" int mrf_ra_write(const mrf_ra_t *a, const char *path)
{
    // src/ra.cu:59-92 layout; the data loop writes every byte for any size
    // (the reference's fixed bufsize over-reads past 2 GiB).
    if (!a || !path || (!a->dims && a->ndims)) return mrf_fail(MRF_E_ARG, "mrf_ra_write: null argument");
    int fd = ::open(path, O_WRONLY | O_TRUNC | O_CREAT, 0644);
    if (fd < 0) return mrf_fail(MRF_E_IO, "unable to open %s for writing", path);
    uint64_t h[6] = {kRaMagic, a->flags, a->eltype, a->elbyte, a->size, a->ndims};
    bool ok = write_full(fd, h, sizeof(h)) &&
              write_full(fd, a->dims, a->ndims * sizeof(uint64_t)) &&
              write_full(fd, a->data, a->size);
    ::close(fd);
    if (!ok) return mrf_fail(MRF_E_IO, "%s: write failed", path);
    return MRF_OK;
}

extern "C" void mrf_ra_free(mrf_ra_t *a)
{
    if (!a) return;
    std::free(a->data);
    std::free(a->dims);
    a->data = nullptr;
    a->dims = nullptr;
}
