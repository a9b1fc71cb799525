// Text input of the fit: Gaussian cube files (read_gauss_cube, repeat.py:770-804) and the value
// blocks of VASP LOCPOT files (read_vasp_potential, repeat.py:668-768).
//
// The reference parses every value with Python's float() in one thread.  Here the value block is
// memory-mapped, cut at whitespace into one range per thread, counted, and parsed in place.  Each
// token goes through the exact fast path of Clinger's method: a decimal with at most 19 significant
// digits whose mantissa m fits in 53 bits and whose power of ten e satisfies |e| <= 22 is m * 10^e
// (or m / 10^-e) with both operands exact, so the one IEEE operation rounds correctly and the result
// is bitwise what float() returns.  Every other token (more digits, larger exponents, inf/nan) goes
// to strtod, which glibc rounds correctly as well.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "rp_internal.h"

namespace rp {
namespace {

int fail(int rc, const std::string &msg) {
    set_global_error(msg);
    return rc;
}

inline bool is_space(char c) { return c == ' ' || c == '\n' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

constexpr double kPow10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,  1e8,  1e9,  1e10, 1e11,
                               1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};

// Parses the token [p, q) as float() would.  Returns false for a token float() rejects.
bool parse_token(const char *p, const char *q, double *out) {
    const char *s = p;
    bool neg = false;
    if (s < q && (*s == '+' || *s == '-')) neg = *s++ == '-';
    uint64_t m = 0;
    int digits = 0, exp10 = 0;
    bool any = false, fast = true;
    while (s < q && *s >= '0' && *s <= '9') {
        any = true;
        if (m == 0 && *s == '0') {
            ++s;
            continue;
        }
        if (digits < 19) {
            m = m * 10 + (uint64_t)(*s - '0');
            ++digits;
        } else {
            fast = false;
            ++exp10;
        }
        ++s;
    }
    if (s < q && *s == '.') {
        ++s;
        while (s < q && *s >= '0' && *s <= '9') {
            any = true;
            if (m == 0 && *s == '0') {
                --exp10;
            } else if (digits < 19) {
                m = m * 10 + (uint64_t)(*s - '0');
                ++digits;
                --exp10;
            } else {
                fast = false;
            }
            ++s;
        }
    }
    if (any && s < q && (*s == 'e' || *s == 'E')) {
        const char *e = s + 1;
        bool eneg = false;
        if (e < q && (*e == '+' || *e == '-')) eneg = *e++ == '-';
        if (e < q && *e >= '0' && *e <= '9') {
            int ev = 0;
            while (e < q && *e >= '0' && *e <= '9') {
                if (ev < 100000) ev = ev * 10 + (*e - '0');
                ++e;
            }
            exp10 += eneg ? -ev : ev;
            s = e;
        }
    }
    if (any && s == q && fast && m <= (1ull << 53) && exp10 >= -22 && exp10 <= 22) {
        double v = (double)m;
        v = exp10 >= 0 ? v * kPow10[exp10] : v / kPow10[-exp10];
        *out = neg ? -v : v;
        return true;
    }
    // general case: strtod on a terminated copy; reject what float() rejects (hex forms, junk)
    char buf[128];
    const size_t len = (size_t)(q - p);
    if (len == 0 || len >= sizeof(buf)) return false;
    std::memcpy(buf, p, len);
    buf[len] = '\0';
    for (size_t i = 0; i < len; ++i)
        if (buf[i] == 'x' || buf[i] == 'X' || buf[i] == 'p' || buf[i] == 'P') return false;
    char *end = nullptr;
    errno = 0;
    const double v = std::strtod(buf, &end);
    if (end != buf + len) return false;
    *out = v;
    return true;
}

struct Range {
    const char *b, *e;
    int64_t count = 0;
    int64_t first = 0;
    bool ok = true;
};

// Counts (out == nullptr) or parses the whitespace-separated tokens of r; parsed values land at
// out[first + k] for first + k < cap.
void scan(Range &r, double *out, int64_t cap) {
    const char *s = r.b;
    int64_t k = 0;
    while (s < r.e) {
        while (s < r.e && is_space(*s)) ++s;
        if (s >= r.e) break;
        const char *t = s;
        while (t < r.e && !is_space(*t)) ++t;
        if (out) {
            double v;
            if (!parse_token(s, t, &v)) {
                r.ok = false;
                return;
            }
            const int64_t at = r.first + k;
            if (at < cap) out[at] = v;
        }
        ++k;
        s = t;
    }
    r.count = k;
}

}  // namespace

// Parses all whitespace-separated values of text[0, len) into out (the first cap of them); returns
// the total token count in *n_found.  Fails on a token float() would reject.
int parse_values(const char *text, int64_t len, double *out, int64_t cap, int n_threads, int64_t *n_found,
                 std::string &err) {
    if (n_threads <= 0) n_threads = (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    const int64_t min_bytes = 1 << 20;
    n_threads = (int)std::max<int64_t>(1, std::min<int64_t>(n_threads, len / min_bytes + 1));
    std::vector<Range> rs(n_threads);
    const char *end = text + len;
    const char *p = text;
    for (int i = 0; i < n_threads; ++i) {
        rs[i].b = p;
        const char *q = i + 1 == n_threads ? end : std::min(end, text + (len * (i + 1)) / n_threads);
        if (q < p) q = p;
        while (q < end && !is_space(*q)) ++q;   // cut at whitespace so no token is split
        rs[i].e = q;
        p = q;
    }
    auto run = [&](double *dst) {
        std::vector<std::thread> th;
        for (int i = 1; i < n_threads; ++i) th.emplace_back([&, i] { scan(rs[i], dst, cap); });
        scan(rs[0], dst, cap);
        for (auto &t : th) t.join();
    };
    run(nullptr);
    int64_t total = 0;
    for (auto &r : rs) {
        r.first = total;
        total += r.count;
    }
    run(out);
    for (auto &r : rs)
        if (!r.ok) {
            err = "malformed value in the data block";
            return RP_EINVAL;
        }
    *n_found = total;
    return RP_OK;
}

namespace {

struct Mapped {
    const char *p = nullptr;
    size_t n = 0;
    int fd = -1;
    ~Mapped() {
        if (p && n) munmap(const_cast<char *>(p), n);
        if (fd >= 0) close(fd);
    }
    bool open(const char *path, std::string &err) {
        fd = ::open(path, O_RDONLY);
        if (fd < 0) {
            err = std::string("cannot open ") + path;
            return false;
        }
        struct stat st;
        if (fstat(fd, &st) != 0) {
            err = std::string("cannot stat ") + path;
            return false;
        }
        n = (size_t)st.st_size;
        if (n == 0) {
            err = std::string("empty file ") + path;
            return false;
        }
        void *m = mmap(nullptr, n, PROT_READ, MAP_PRIVATE, fd, 0);
        if (m == MAP_FAILED) {
            n = 0;
            err = std::string("cannot map ") + path;
            return false;
        }
        p = static_cast<const char *>(m);
        return true;
    }
};

// Line cursor over the mapped file.
struct Lines {
    const char *p, *end;
    bool next(std::string &line) {
        if (p >= end) return false;
        const char *q = static_cast<const char *>(std::memchr(p, '\n', (size_t)(end - p)));
        if (!q) q = end;
        line.assign(p, q);
        p = q < end ? q + 1 : end;
        return true;
    }
};

// Splits a header line into tokens.
std::vector<std::string> fields(const std::string &line) {
    std::vector<std::string> f;
    size_t i = 0;
    while (i < line.size()) {
        while (i < line.size() && is_space(line[i])) ++i;
        size_t j = i;
        while (j < line.size() && !is_space(line[j])) ++j;
        if (j > i) f.emplace_back(line.substr(i, j - i));
        i = j;
    }
    return f;
}

bool to_int(const std::string &s, long &v) {
    char *e = nullptr;
    errno = 0;
    v = std::strtol(s.c_str(), &e, 10);
    return e && *e == '\0' && errno == 0;
}

bool to_double(const std::string &s, double &v) { return parse_token(s.data(), s.data() + s.size(), &v); }

// The header: two comment lines, "n_atoms ox oy oz", three "n_i ax ay az" lines, then n_atoms lines
// "Z charge x y z".  Returns the byte offset of the values.
int cube_header(const char *text, size_t n, rp_cube_header *h, std::vector<int32_t> *Z, std::vector<double> *xyz,
                std::string &err) {
    Lines ln{text, text + n};
    std::string line;
    for (int i = 0; i < 2; ++i)
        if (!ln.next(line)) {
            err = "cube file ends inside the comment lines";
            return RP_EINVAL;
        }
    if (!ln.next(line)) {
        err = "cube file has no atom-count line";
        return RP_EINVAL;
    }
    auto f = fields(line);
    long na = 0;
    if (f.empty() || !to_int(f[0], na)) {
        err = "bad atom-count line";
        return RP_EINVAL;
    }
    if (na < 0) {
        err = "cube files with a negative atom count are not supported";
        return RP_EINVAL;
    }
    h->n_atoms = (int32_t)na;
    for (int c = 0; c < 3; ++c) h->origin[c] = 0.0;
    for (int c = 0; c < 3 && c + 1 < (int)f.size(); ++c) to_double(f[c + 1], h->origin[c]);
    for (int i = 0; i < 3; ++i) {
        if (!ln.next(line)) {
            err = "cube file ends inside the grid lines";
            return RP_EINVAL;
        }
        f = fields(line);
        long ng = 0;
        if (f.size() < 4 || !to_int(f[0], ng) || ng <= 0) {
            err = "bad grid line";
            return RP_EINVAL;
        }
        h->n_grid[i] = (int32_t)ng;
        for (int c = 0; c < 3; ++c)
            if (!to_double(f[c + 1], h->axis[3 * i + c])) {
                err = "bad grid vector";
                return RP_EINVAL;
            }
    }
    if (Z) Z->resize((size_t)na);
    if (xyz) xyz->resize((size_t)na * 3);
    for (long a = 0; a < na; ++a) {
        if (!ln.next(line)) {
            err = "cube file ends inside the atom lines";
            return RP_EINVAL;
        }
        f = fields(line);
        long z = 0;
        double x[3];
        if (f.size() < 5 || !to_int(f[0], z) || !to_double(f[2], x[0]) || !to_double(f[3], x[1]) ||
            !to_double(f[4], x[2])) {
            err = "bad atom line " + std::to_string(a + 1);
            return RP_EINVAL;
        }
        if (Z) (*Z)[a] = (int32_t)z;
        if (xyz)
            for (int c = 0; c < 3; ++c) (*xyz)[3 * a + c] = x[c];
    }
    h->values_offset = (int64_t)(ln.p - text);
    return RP_OK;
}

}  // namespace

}  // namespace rp

using namespace rp;

extern "C" {

int rp_cube_read_header(const char *path, rp_cube_header *h) {
    if (!path || !h) return fail(RP_EINVAL, "null argument");
    Mapped m;
    std::string err;
    if (!m.open(path, err)) return fail(RP_EINVAL, err);
    std::memset(h, 0, sizeof(*h));
    const int rc = cube_header(m.p, m.n, h, nullptr, nullptr, err);
    if (rc != RP_OK) return fail(rc, err);
    return RP_OK;
}

int rp_cube_read(const char *path, const rp_cube_header *h, int32_t *atomic_number, double *atom_xyz, double *esp,
                 int n_threads) {
    if (!path || !h || !esp) return fail(RP_EINVAL, "null argument");
    Mapped m;
    std::string err;
    if (!m.open(path, err)) return fail(RP_EINVAL, err);
    rp_cube_header hh{};
    std::vector<int32_t> Z;
    std::vector<double> xyz;
    int rc = cube_header(m.p, m.n, &hh, &Z, &xyz, err);
    if (rc != RP_OK) return fail(rc, err);
    if (hh.n_atoms != h->n_atoms || hh.n_grid[0] != h->n_grid[0] || hh.n_grid[1] != h->n_grid[1] ||
        hh.n_grid[2] != h->n_grid[2])
        return fail(RP_EINVAL, "header does not match the file");
    if (atomic_number) std::copy(Z.begin(), Z.end(), atomic_number);
    if (atom_xyz) std::copy(xyz.begin(), xyz.end(), atom_xyz);
    const int64_t n_total = (int64_t)hh.n_grid[0] * hh.n_grid[1] * hh.n_grid[2];
    int64_t found = 0;
    rc = parse_values(m.p + hh.values_offset, (int64_t)m.n - hh.values_offset, esp, n_total, n_threads, &found, err);
    if (rc != RP_OK) return fail(rc, err);
    if (found < n_total)
        return fail(RP_EINVAL, "cube holds " + std::to_string(found) + " values, grid needs " +
                                               std::to_string(n_total));
    return RP_OK;
}

int rp_parse_values(const char *text, int64_t len, double *out, int64_t cap, int n_threads, int64_t *n_found) {
    if (!text || len < 0 || (!out && cap > 0) || !n_found) return fail(RP_EINVAL, "null argument");
    std::string err;
    const int rc = parse_values(text, len, out, cap, n_threads, n_found, err);
    if (rc != RP_OK) return fail(rc, err);
    return RP_OK;
}

}  // extern "C"
