// Text files of a run in the reference's CSV layout (constants, simData, energy):
// formatting, parsing and the last-line reader a continued run starts from.  Plain
// C++ host code.  Numbers are printed as printf's "%.*g" prints them, which is what C++
// streams use for a double in the default float field: std::to_chars in the general
// format with a precision is specified to give exactly that text, about three times
// faster than snprintf.
#include "spacelab/sl_csv.h"

#include "spacelab/sl_api.h"

#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

namespace {

void put(std::string& s, double v, int digits)
{
    char b[64];
    const std::to_chars_result r = std::to_chars(b, b + sizeof b, v, std::chars_format::general, digits);
    s.append(b, (size_t)(r.ptr - b));
}

long long emit(const std::string& s, char* buf, long long cap)
{
    if (buf && cap > 0) {
        const size_t k = s.size() < (size_t)(cap - 1) ? s.size() : (size_t)(cap - 1);
        std::memcpy(buf, s.data(), k);
        buf[k] = '\0';
    }
    return (long long)s.size();
}

// |v| with the reference's operation order (x*x + y*y) + z*z
double norm3(const double* v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

// strtod over one comma- or newline-terminated field; advances *p past the separator
bool field(const char*& p, double* out)
{
    if (!*p) return false;
    char* end = nullptr;
    const double v = std::strtod(p, &end);
    if (end == p) return false;
    if (*end && *end != ',' && *end != '\n' && *end != '\r') return false;
    if (out) *out = v;
    p = *end ? end + 1 : end;
    return true;
}

// skips one field whatever it holds
bool skip_field(const char*& p)
{
    if (!*p) return false;
    while (*p && *p != ',' && *p != '\n') p++;
    if (*p) p++;
    return true;
}

}  // namespace

extern "C" {

long long sl_csv_sim_header(int n, char* buf, long long cap)
{
    if (n < 1) return SL_E_ARG;
    static const char* cols[SL_CSV_PROPERTIES] = {"x", "y", "z", "wx", "wy", "wz", "wmag", "vx", "vy", "vz", "bound"};
    std::string s;
    s.reserve((size_t)n * 64);
    for (int b = 0; b < n; b++)
        for (int c = 0; c < SL_CSV_PROPERTIES; c++) {
            if (b || c) s += ',';
            s += cols[c];
            s += std::to_string(b);
        }
    return emit(s, buf, cap);
}

long long sl_csv_sim_row(int n, const double* pos3, const double* w3, const double* vel3, char* buf, long long cap)
{
    if (n < 1 || !pos3 || !w3 || !vel3) return SL_E_ARG;
    std::string s;
    s.reserve((size_t)n * 128);
    const int d = SL_CSV_STATE_DIGITS;
    for (int b = 0; b < n; b++) {
        if (b) s += ',';
        const double* x = pos3 + 3 * (size_t)b;
        const double* w = w3 + 3 * (size_t)b;
        const double* v = vel3 + 3 * (size_t)b;
        for (int k = 0; k < 3; k++) {
            put(s, x[k], d);
            s += ',';
        }
        for (int k = 0; k < 3; k++) {
            put(s, w[k], d);
            s += ',';
        }
        put(s, norm3(w), d);
        for (int k = 0; k < 3; k++) {
            s += ',';
            put(s, v[k], d);
        }
        s += ",0";
    }
    return emit(s, buf, cap);
}

long long sl_csv_energy_row(double time, double PE, double KE, const double* p3, const double* L3, int digits,
                            char* buf, long long cap)
{
    if (!p3 || !L3 || digits < 1 || digits > 17) return SL_E_ARG;
    std::string s;
    put(s, time, digits);
    s += ',';
    put(s, PE, digits);
    s += ',';
    put(s, KE, digits);
    s += ',';
    put(s, PE + KE, digits);
    s += ',';
    put(s, norm3(p3), digits);
    s += ',';
    put(s, norm3(L3), digits);
    return emit(s, buf, cap);
}

long long sl_csv_constants(int n, const double* R, const double* m, const double* moi, char* buf, long long cap)
{
    if (n < 1 || !R || !m || !moi) return SL_E_ARG;
    std::string s;
    s.reserve((size_t)n * 40);
    for (int b = 0; b < n; b++) {
        put(s, R[b], SL_CSV_STATE_DIGITS);
        s += ',';
        put(s, m[b], SL_CSV_STATE_DIGITS);
        s += ',';
        put(s, moi[b], SL_CSV_STATE_DIGITS);
        s += '\n';
    }
    return emit(s, buf, cap);
}

int sl_csv_count_balls(const char* row)
{
    if (!row) return SL_E_ARG;
    long long commas = 0;
    for (const char* p = row; *p && *p != '\n'; p++) commas += *p == ',';
    return (int)(commas / SL_CSV_PROPERTIES + 1);
}

int sl_csv_parse_sim_row(const char* row, int n, double* pos3, double* w3, double* vel3)
{
    if (!row || n < 1 || !pos3 || !w3 || !vel3) return SL_E_ARG;
    const char* p = row;
    for (int b = 0; b < n; b++) {
        for (int k = 0; k < 3; k++)
            if (!field(p, pos3 + 3 * (size_t)b + k)) return SL_E_ARG;
        for (int k = 0; k < 3; k++)
            if (!field(p, w3 + 3 * (size_t)b + k)) return SL_E_ARG;
        if (!skip_field(p)) return SL_E_ARG;   // |w|
        for (int k = 0; k < 3; k++)
            if (!field(p, vel3 + 3 * (size_t)b + k)) return SL_E_ARG;
        // the remaining columns of the ball (bound); the last ball's may end the row
        for (int k = 10; k < SL_CSV_PROPERTIES; k++)
            if (!skip_field(p) && !(b == n - 1 && k == SL_CSV_PROPERTIES - 1)) return SL_E_ARG;
    }
    return SL_OK;
}

int sl_csv_parse_constants(const char* text, int n, double* R, double* m, double* moi)
{
    if (!text || n < 1 || !R || !m || !moi) return SL_E_ARG;
    const char* p = text;
    for (int b = 0; b < n; b++) {
        // one line per ball; columns after the third are ignored, as loadConsts reads
        // three fields of each line
        const char* eol = p;
        while (*eol && *eol != '\n') eol++;
        const std::string line(p, eol);
        const char* q = line.c_str();
        if (!field(q, R + b) || !field(q, m + b) || !field(q, moi + b)) return SL_E_ARG;
        p = *eol ? eol + 1 : eol;
    }
    return SL_OK;
}

int sl_csv_energy_columns(const char* header)
{
    if (!header) return SL_E_ARG;
    size_t len = 0;
    while (header[len] && header[len] != '\n' && header[len] != '\r') len++;
    const std::string h(header, len);
    if (h == SL_CSV_ENERGY_HEADER) return 6;
    if (h == SL_CSV_ENERGY_HEADER_BOUND) return 9;
    return SL_E_ARG;
}

int sl_csv_parse_energy_row(const char* row, int ncols, double* out)
{
    if (!row || !out || (ncols != 6 && ncols != 9)) return SL_E_ARG;
    const char* p = row;
    for (int k = 0; k < ncols; k++) {
        if (!field(p, out + k)) return SL_E_ARG;
        // the row ends after its last column
        if (k + 1 < ncols && (p == row || p[-1] != ',')) return SL_E_ARG;
    }
    while (*p == '\r' || *p == '\n') p++;
    return *p ? SL_E_ARG : SL_OK;
}

long long sl_csv_last_line(const char* path, char* buf, long long cap)
{
    if (!path) return SL_E_ARG;
    FILE* f = std::fopen(path, "rb");
    if (!f) return SL_E_ARG;
    // scan back from the end in blocks for the final '\n'
    if (std::fseek(f, 0, SEEK_END) != 0) {
        std::fclose(f);
        return SL_E_ARG;
    }
    const long size = std::ftell(f);
    long start = 0;
    std::string tail;
    char blk[4096];
    long pos = size;
    bool found = false;
    while (pos > 0 && !found) {
        const long len = pos >= (long)sizeof blk ? (long)sizeof blk : pos;
        pos -= len;
        std::fseek(f, pos, SEEK_SET);
        if (std::fread(blk, 1, (size_t)len, f) != (size_t)len) {
            std::fclose(f);
            return SL_E_ARG;
        }
        for (long i = len - 1; i >= 0; i--)
            if (blk[i] == '\n') {
                start = pos + i + 1;
                found = true;
                break;
            }
    }
    // getLastLine stops its backward scan at the second byte: when the only '\n' left is
    // the first byte, it reads the first line, which is empty
    if (found && start == 1) start = size;
    std::string line((size_t)(size - start), '\0');
    std::fseek(f, start, SEEK_SET);
    const size_t got = line.empty() ? 0 : std::fread(&line[0], 1, line.size(), f);
    std::fclose(f);
    if (got != line.size()) return SL_E_ARG;
    return emit(line, buf, cap);
}

}  // extern "C"
