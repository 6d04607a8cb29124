// fgl_ingest.hip -- mesh ingest into a resident device mesh: LoadSTL
// (stl.go:23-154) and LoadOBJ (obj.go:19-79).
//
// Binary STL goes to the device as the file's bytes: one thread per 50-byte
// record widens the three float32 vertices to float64 and writes the face
// normal (Triangle.Normal, triangle.go:33-37) to all three vertices, which
// is what loadSTLB stores.  ASCII STL and OBJ are text: they are tokenised
// on the host (strconv.ParseFloat / ParseInt semantics), uploaded as SoA
// doubles, and FixNormals (triangle.go:46-58) runs on the device.
#include "fgl_internal.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>

namespace fgl {

// Triangle.Normal: Cross (vector.go:76-81) of the two edges, Normalize
FGL_D void face_normal(const double *p, double n[3]) {
    const double e1x = p[3] - p[0], e1y = p[4] - p[1], e1z = p[5] - p[2];
    const double e2x = p[6] - p[0], e2y = p[7] - p[1], e2z = p[8] - p[2];
    const double cx = e1y * e2z - e1z * e2y, cy = e1z * e2x - e1x * e2z, cz = e1x * e2y - e1y * e2x;
    const double r = 1.0 / sqrt(cx * cx + cy * cy + cz * cz);
    n[0] = cx * r;
    n[1] = cy * r;
    n[2] = cz * r;
}

// little-endian float32 at a 2-byte-aligned address, widened
FGL_D double le_f32(const uint16_t *h) {
    return (double)__uint_as_float((uint32_t)h[0] | ((uint32_t)h[1] << 16));
}

// records start at byte 84 + 50 i: vertices at +12, +24, +36
__global__ __launch_bounds__(256) void k_stl_decode(const uint8_t *file, int64_t count, double *pos, double *nrm) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
        const uint16_t *h = reinterpret_cast<const uint16_t *>(file + 84 + 50 * i + 12);
        double p[9];
#pragma unroll
        for (int k = 0; k < 9; k++) p[k] = le_f32(h + 2 * k);
        double n[3];
        face_normal(p, n);
#pragma unroll
        for (int k = 0; k < 9; k++) pos[9 * i + k] = p[k];
#pragma unroll
        for (int v = 0; v < 3; v++)
#pragma unroll
            for (int k = 0; k < 3; k++) nrm[9 * i + 3 * v + k] = n[k];
    }
}

// FixNormals: a vertex normal equal to the zero Vector becomes the face normal
__global__ __launch_bounds__(256) void k_fix_normals(const double *pos, double *nrm, int64_t ntri) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ntri; i += (int64_t)gridDim.x * blockDim.x) {
        double p[9];
#pragma unroll
        for (int k = 0; k < 9; k++) p[k] = pos[9 * i + k];
        double n[3];
        face_normal(p, n);
#pragma unroll
        for (int v = 0; v < 3; v++) {
            double *q = nrm + 9 * i + 3 * v;
            if (q[0] == 0.0 && q[1] == 0.0 && q[2] == 0.0) {
                q[0] = n[0];
                q[1] = n[1];
                q[2] = n[2];
            }
        }
    }
}

static unsigned ingest_grid(int64_t n) {
    const int64_t b = (n + 255) / 256;
    return (unsigned)(b < 1 ? 1 : (b > 8192 ? 8192 : b));
}

hipError_t launch_stl_decode(const uint8_t *file, int64_t count, double *pos, double *nrm, hipStream_t s) {
    hipLaunchKernelGGL(k_stl_decode, dim3(ingest_grid(count)), dim3(256), 0, s, file, count, pos, nrm);
    return hipGetLastError();
}
hipError_t launch_fix_normals(const double *pos, double *nrm, int64_t ntri, hipStream_t s) {
    hipLaunchKernelGGL(k_fix_normals, dim3(ingest_grid(ntri)), dim3(256), 0, s, pos, nrm, ntri);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// host-side tokenising (text formats)

namespace {

bool ws(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r'; }

bool ieq(const char *a, size_t n, const char *lit) {
    size_t m = strlen(lit);
    if (n != m) return false;
    for (size_t i = 0; i < n; i++) {
        char c = a[i];
        if (c >= 'A' && c <= 'Z') c = (char)(c - 'A' + 'a');
        if (c != lit[i]) return false;
    }
    return true;
}

}  // namespace

// strconv.ParseFloat(s, 64), a malformed token giving 0: decimal
// [+-](digits[.digits] | .digits)([eE][+-]digits), hexadecimal
// [+-]0x(hex[.hex] | .hex)p[+-]digits, or [+-]inf / infinity / nan (any
// case); the value is the correctly rounded double (strtod).  Underscores
// (which Go allows inside a hexadecimal literal) read as malformed.
double ingest_parse_float(const char *s, size_t n) {
    if (n == 0) return 0.0;
    size_t i = 0;
    bool neg = false;
    if (s[0] == '+' || s[0] == '-') {
        neg = s[0] == '-';
        i = 1;
    }
    if (ieq(s + i, n - i, "inf") || ieq(s + i, n - i, "infinity")) return neg ? -HUGE_VAL : HUGE_VAL;
    if (ieq(s + i, n - i, "nan")) return NAN;
    if (n - i > 2 && s[i] == '0' && (s[i + 1] == 'x' || s[i + 1] == 'X')) {
        // hexadecimal mantissa, binary exponent required (p / P)
        auto hex = [](char c) { return (c >= '0' && c <= '9') || (c >= 'a' && c <= 'f') || (c >= 'A' && c <= 'F'); };
        size_t j = i + 2, digits = 0;
        while (j < n && hex(s[j])) ++j, ++digits;
        if (j < n && s[j] == '.')
            for (++j; j < n && hex(s[j]); ++j) ++digits;
        if (digits == 0 || j >= n || (s[j] != 'p' && s[j] != 'P')) return 0.0;
        size_t k = j + 1, ed = 0;
        if (k < n && (s[k] == '+' || s[k] == '-')) ++k;
        while (k < n && s[k] >= '0' && s[k] <= '9') ++k, ++ed;
        if (ed == 0 || k != n) return 0.0;
        std::string buf(s, n);
        return strtod(buf.c_str(), nullptr);
    }
    size_t j = i, digits = 0;
    while (j < n && s[j] >= '0' && s[j] <= '9') ++j, ++digits;
    if (j < n && s[j] == '.')
        for (++j; j < n && s[j] >= '0' && s[j] <= '9'; ++j) ++digits;
    if (digits == 0) return 0.0;
    if (j < n && (s[j] == 'e' || s[j] == 'E')) {
        size_t k = j + 1, ed = 0;
        if (k < n && (s[k] == '+' || s[k] == '-')) ++k;
        while (k < n && s[k] >= '0' && s[k] <= '9') ++k, ++ed;
        if (ed == 0) return 0.0;
        j = k;
    }
    if (j != n) return 0.0;
    std::string buf(s, n);
    return strtod(buf.c_str(), nullptr);
}

// strconv.ParseInt(s, 0, 0), an error giving 0 and an out-of-range value
// the clamped one: sign, 0x / 0o / 0b / leading-zero octal, underscores
// between digits or right after a prefix
int64_t ingest_parse_int(const char *s, size_t n) {
    size_t i = 0;
    bool neg = false;
    if (i < n && (s[i] == '+' || s[i] == '-')) neg = s[i++] == '-';
    if (i >= n) return 0;
    int base = 10;
    bool after_prefix = false;
    if (s[i] == '0' && i + 1 < n) {
        const char c = s[i + 1];
        if (c == 'x' || c == 'X') base = 16, i += 2;
        else if (c == 'o' || c == 'O') base = 8, i += 2;
        else if (c == 'b' || c == 'B') base = 2, i += 2;
        else base = 8, i += 1;
        after_prefix = true;
    }
    uint64_t v = 0;
    bool ok_us = after_prefix;  // an underscore may follow a digit or the prefix
    int digits = 0;
    bool last_us = false;
    for (; i < n; i++) {
        const char c = s[i];
        if (c == '_') {
            if (!ok_us) return 0;
            ok_us = false;
            last_us = true;
            continue;
        }
        int d = -1;
        if (c >= '0' && c <= '9') d = c - '0';
        else if (c >= 'a' && c <= 'z') d = c - 'a' + 10;
        else if (c >= 'A' && c <= 'Z') d = c - 'A' + 10;
        if (d < 0 || d >= base) return 0;
        v = v > (UINT64_MAX - (uint64_t)d) / (uint64_t)base ? UINT64_MAX : v * (uint64_t)base + (uint64_t)d;
        ++digits;
        ok_us = true;
        last_us = false;
    }
    if (last_us) return 0;
    if (digits == 0 && base != 8) return 0;
    if (!neg) return v > (uint64_t)INT64_MAX ? INT64_MAX : (int64_t)v;
    return v > (uint64_t)INT64_MAX + 1 ? INT64_MIN : (int64_t)(0 - v);
}

// strings.Fields over one line
void ingest_fields(const char *a, const char *b, std::vector<std::pair<const char *, size_t>> &out) {
    out.clear();
    const char *x = a;
    while (x < b) {
        while (x < b && ws(*x)) ++x;
        if (x >= b) break;
        const char *y = x;
        while (y < b && !ws(*y)) ++y;
        out.emplace_back(x, (size_t)(y - x));
        x = y;
    }
}

}  // namespace fgl
