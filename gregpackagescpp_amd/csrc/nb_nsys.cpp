// nb_nsys.cpp -- .nsys snapshot reader and writer (host code).
//
// Reference: the 96-byte nsys_file_header and nsys_file_chunk
// (gregsys.hpp:1444-1477), system::to_nsys (:3214-3258) and load_from_nsys
// (:943-1400) with its checks and messages, check_dt (:1401-1405) and the
// G-from-ratios arithmetic shared with parse_units_format (:1281-1284).
//
// The reference fixes the stored types through its template parameters; this
// reader accepts every type the header can describe (floating M/R/T of 4, 8 or
// 16 bytes, integral of 1, 2, 4 or 8 bytes) and converts to FP64.  Checks that
// the reference makes against its own template types are left to the caller
// (the C++ facade repeats them with the reference's messages).
#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "nb_system.hpp"

using namespace nbe;

namespace {

constexpr size_t kHeaderSize = 96;
constexpr int kLdSize = 16;  // sizeof(long double), x87 80-bit in a 16-byte slot

struct NsysError {
    int code;
    std::string msg;
};

[[noreturn]] void load_error(const std::string &m) { throw NsysError{NB_ERR_NSYS_LOAD, m}; }
[[noreturn]] void format_error(const std::string &m) { throw NsysError{NB_ERR_NSYS_FORMAT, m}; }

// A stored scalar type: floating (4, 8, 16 bytes) or integral (1, 2, 4, 8).
struct Field {
    bool fp = true;
    bool is_signed = false;
    int size = 8;
    int align() const { return size; }
    long double get(const unsigned char *p) const {
        if (fp) {
            if (size == 4) {
                float f;
                std::memcpy(&f, p, 4);
                return f;
            }
            if (size == 8) {
                double d;
                std::memcpy(&d, p, 8);
                return d;
            }
            long double l = 0;
            std::memcpy(&l, p, 10);  // x87 extended: 64-bit significand, sign and exponent
            return l;
        }
        uint64_t u = 0;
        std::memcpy(&u, p, (size_t)size);
        if (is_signed && size < 8 && (u >> (8 * size - 1)) & 1) u |= ~uint64_t(0) << (8 * size);
        return is_signed ? (long double)(int64_t)u : (long double)u;
    }
};

Field field_of(int d_types, int fp_bit, int signed_bit, int size, const char *what) {
    Field f;
    f.fp = (d_types & fp_bit) != 0;
    f.is_signed = (d_types & signed_bit) != 0;
    f.size = size;
    const bool ok = f.fp ? (size == 4 || size == 8 || size == kLdSize)
                         : (size == 1 || size == 2 || size == 4 || size == 8);
    if (!ok) {
        char msg[160];
        snprintf(msg, sizeof msg,
                 "Size mismatch error: %s values are stored as %s of %d bytes, which no type of this engine "
                 "matches.\n",
                 what, f.fp ? "floating point" : "integers", size);
        load_error(msg);
    }
    return f;
}

struct Layout {
    Field m, r, t;
    size_t off_m = 0, off_r = 0, off_t = 0, chunk = 0;
};

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// nsys_file_chunk: char r_coeff[16], then M, R and six T in natural alignment.
Layout layout_of(const Field &m, const Field &r, const Field &t) {
    Layout L{m, r, t};
    size_t off = 16;
    off = align_up(off, (size_t)m.align());
    L.off_m = off;
    off += (size_t)m.size;
    off = align_up(off, (size_t)r.align());
    L.off_r = off;
    off += (size_t)r.size;
    off = align_up(off, (size_t)t.align());
    L.off_t = off;
    off += 6 * (size_t)t.size;
    const size_t max_align = std::max<size_t>({1, (size_t)m.align(), (size_t)r.align(), (size_t)t.align()});
    L.chunk = align_up(off, max_align);
    return L;
}

// nsys_file_size (:943-946): header + chunks, padded to a multiple of 4.
uint64_t file_size_for(uint64_t n, size_t chunk) {
    const uint64_t body = n * (uint64_t)chunk;
    const unsigned rem = (unsigned)(body % 4);
    return kHeaderSize + body + (rem ? 4 - rem : 0);
}

struct Parsed {
    int d_types = 0, time_size = 0, m_size = 0, r_size = 0, t_size = 0, rest_size = 0;
    uint64_t g[6] = {1, 1, 1, 1, 1, 1};
    uint64_t n = 0;
    long double time = 0, dt = 1;
    std::vector<double> m, r, pos, vel, rest;
};

long double ld_time_field(const unsigned char *p, int time_size, bool fp) {
    if (fp) {
        if (time_size == kLdSize) {
            long double l = 0;
            std::memcpy(&l, p, 10);
            return l;
        }
        if (time_size == 8) {
            double d;
            std::memcpy(&d, p, 8);
            return d;
        }
        float f;
        std::memcpy(&f, p, 4);
        return f;
    }
    uint64_t u = 0;
    std::memcpy(&u, p, (size_t)time_size);
    return (long double)u;
}

// load_from_nsys (:948-1400): same checks, in the same order, with the same
// messages where the reference's checks apply to an untyped reader.
void read_nsys(const char *path, bool check_nan_inf, bool want_bodies, Parsed &P) {
    struct stat st {};
    if (stat(path, &st) == -1) load_error("Error obtaining .nsys file information. No data loaded.\n");
    if (!S_ISREG(st.st_mode)) load_error(".nsys file provided is not a regular file. Could not load data.\n");
    if (st.st_size < (off_t)kHeaderSize) load_error("Insufficient .nsys file size. No data loaded.\n");
    std::ifstream in(path, std::ios_base::in | std::ios_base::binary);
    if (!in) load_error("Error opening .nsys file. Could not load data.\n");
    unsigned char h[kHeaderSize];
    in.read((char *)h, kHeaderSize);
    // the reference's signature test rejects a file only when both bytes differ
    if (h[0] != 'N' && h[1] != 'S') format_error("Invalid .nsys format: invalid file signature.\n");
    P.d_types = h[2];
    P.time_size = h[3];
    if ((P.d_types & 0b1000000) != 0)  // the reference tests bit 6 here (the signed-T flag)
        format_error("Invalid .nsys format: msb occupied in d_types field (should be zero).\n");
    P.m_size = h[36];
    P.r_size = h[37];
    P.t_size = h[38];
    P.rest_size = h[39];
    const Field fm = field_of(P.d_types, 0b10, 0b10000, P.m_size, "mass");
    const Field fr = field_of(P.d_types, 0b100, 0b100000, P.r_size, "radius");
    const Field ft = field_of(P.d_types, 0b1000, 0b1000000, P.t_size, "position and velocity");
    if (P.rest_size > kLdSize) {
        char msg[160];
        snprintf(msg, sizeof msg,
                 "Size mismatch error: coefficient of restitution data size (%d bytes) larger than widest floating "
                 "point data type available (long double = %d bytes).\n",
                 P.rest_size, kLdSize);
        format_error(msg);
    }
    if (P.rest_size != kLdSize && P.rest_size != 8 && P.rest_size != 4) {
        char msg[200];
        snprintf(msg, sizeof msg,
                 "Size mismatch error: coefficient of restitution data size (%d bytes) does not match any floating "
                 "point data type available:\nfloat = 4 bytes,\ndouble = 8 bytes,\nlong double = 16 bytes.\n",
                 P.rest_size);
        format_error(msg);
    }
    std::memcpy(P.g, h + 40, 6 * 8);
    std::memcpy(&P.n, h + 88, 8);
    const Layout L = layout_of(fm, fr, ft);
    if ((uint64_t)st.st_size != file_size_for(P.n, L.chunk)) {
        char msg[200];
        snprintf(msg, sizeof msg,
                 "Error: file size does not match expected file size based on number of bodies.\nExpected: %llu "
                 "bytes. Instead got: %zu bytes. No data loaded.\n",
                 (unsigned long long)file_size_for(P.n, L.chunk), (size_t)st.st_size);
        format_error(msg);
    }
    const bool fp_time = (P.d_types & 1) != 0;
    if (fp_time) {
        if (P.time_size != kLdSize && P.time_size != 8 && P.time_size != 4) {
            char msg[220];
            snprintf(msg, sizeof msg,
                     "Size mismatch error: reported size of data type for time values in .nsys file (%d bytes) does "
                     "not match any floating point data type available:\nfloat = 4 bytes,\ndouble = 8 bytes,\n"
                     "long double = 16 bytes.\n",
                     P.time_size);
            format_error(msg);
        }
    } else if (P.time_size > 8) {
        char msg[220];
        snprintf(msg, sizeof msg,
                 "Size mismatch error: reported size of data type for time values in .nsys file (%d bytes) larger "
                 "than largest unsigned integral data types available:\nuint64_t = 8 bytes,\nunsigned long long = "
                 "8 bytes\n",
                 P.time_size);
        format_error(msg);
    }
    P.time = ld_time_field(h + 4, P.time_size, fp_time);
    P.dt = ld_time_field(h + 20, P.time_size, fp_time);
    if (check_nan_inf && (std::isnan(P.dt) || std::isinf(P.dt))) format_error("Error: time-step is NaN or INF.");
    if (P.dt <= 0) throw ApiError(NB_ERR_INVALID_ARGUMENT, "Time-step cannot be zero or negative.\n");
    if (!want_bodies || P.n == 0) return;

    std::vector<unsigned char> buf(L.chunk * (size_t)P.n);
    in.read((char *)buf.data(), (std::streamsize)buf.size());
    const size_t n = (size_t)P.n;
    P.m.resize(n);
    P.r.resize(n);
    P.rest.resize(n);
    P.pos.resize(3 * n);
    P.vel.resize(3 * n);
    const Field frest{true, false, P.rest_size};
    static const char *names[6] = {"xpos", "ypos", "zpos", "xvel", "yvel", "zvel"};
    for (size_t i = 0; i < n; ++i) {
        const unsigned char *c = buf.data() + i * L.chunk;
        auto bad = [&](const char *field) {
            char msg[128];
            snprintf(msg, sizeof msg, "Error: NaN or INF encountered in chunk with index %llu. Invalid field: \"%s\"\n",
                     (unsigned long long)i, field);
            load_error(msg);
        };
        const long double rc = frest.get(c);
        const long double mv = fm.get(c + L.off_m);
        const long double rv = fr.get(c + L.off_r);
        long double tv[6];
        for (int k = 0; k < 6; ++k) tv[k] = ft.get(c + L.off_t + (size_t)k * (size_t)ft.size);
        if (check_nan_inf) {
            if (std::isnan(rc) || std::isinf(rc)) bad("r_coeff");
            if (fm.fp && (std::isnan(mv) || std::isinf(mv))) bad("mass");
            if (fr.fp && (std::isnan(rv) || std::isinf(rv))) bad("radius");
            if (ft.fp)
                for (int k = 0; k < 6; ++k)
                    if (std::isnan(tv[k]) || std::isinf(tv[k])) bad(names[k]);
        }
        P.rest[i] = (double)rc;
        P.m[i] = (double)mv;
        P.r[i] = (double)rv;
        for (int k = 0; k < 3; ++k) {
            P.pos[3 * i + (size_t)k] = (double)tv[k];
            P.vel[3 * i + (size_t)k] = (double)tv[3 + k];
        }
    }
}

void put_ld(unsigned char *p, long double v) {
    std::memset(p, 0, kLdSize);
    std::memcpy(p, &v, 10);
}

// to_nsys (:3214-3258) with M = R = T of value_size bytes.
int64_t write_nsys(const char *path, long double time, long double dt, const uint64_t g[6], int64_t n,
                   const double *m, const double *r, const double *pos, const double *vel, const double *rest,
                   int value_size) {
    if (value_size != 8 && value_size != kLdSize)
        throw ApiError(NB_ERR_INVALID_ARGUMENT, "nb_nsys_write: value_size must be 8 (double) or 16 (long double)");
    if (n < 0 || (n > 0 && (!m || !pos || !vel)))
        throw ApiError(NB_ERR_INVALID_ARGUMENT, "nb_nsys_write: bad arguments");
    unsigned char h[kHeaderSize] = {};
    h[0] = 'N';
    h[1] = 'S';
    h[2] = 1 + (1 << 1) + (1 << 2) + (1 << 3);  // floating time, M, R and T
    h[3] = kLdSize;
    put_ld(h + 4, time);
    put_ld(h + 20, dt);
    h[36] = h[37] = h[38] = (unsigned char)value_size;
    h[39] = kLdSize;
    std::memcpy(h + 40, g, 6 * 8);
    const uint64_t nb = (uint64_t)n;
    std::memcpy(h + 88, &nb, 8);
    const Field f{true, false, value_size};
    const Layout L = layout_of(f, f, f);
    std::ofstream out(path, std::ios_base::trunc | std::ios_base::binary);
    if (!out) throw NsysError{NB_ERR_INVALID_ARGUMENT, std::string("cannot open ") + path + " for writing"};
    out.write((const char *)h, kHeaderSize);
    std::vector<unsigned char> c(L.chunk);
    auto put = [&](size_t off, double v) {
        if (value_size == 8) std::memcpy(c.data() + off, &v, 8);
        else put_ld(c.data() + off, (long double)v);
    };
    for (int64_t i = 0; i < n; ++i) {
        std::fill(c.begin(), c.end(), 0);
        put_ld(c.data(), rest ? (long double)rest[i] : 1.0L);
        put(L.off_m, m[i]);
        put(L.off_r, r ? r[i] : 0.0);
        for (int k = 0; k < 3; ++k) {
            put(L.off_t + (size_t)k * (size_t)value_size, pos[3 * i + k]);
            put(L.off_t + (size_t)(3 + k) * (size_t)value_size, vel[3 * i + k]);
        }
        out.write((const char *)c.data(), (std::streamsize)c.size());
    }
    const uint64_t size = file_size_for(nb, L.chunk);
    const uint64_t written = kHeaderSize + nb * L.chunk;
    if (size > written) out.write("\0\0\0", (std::streamsize)(size - written));
    if (!out) throw NsysError{NB_ERR_INVALID_ARGUMENT, std::string("write to ") + path + " failed"};
    return (int64_t)size;
}

// G = 66743 * (m_num d_num^3 t_num^2) / (m_den d_den^3 t_den^2) / 1e15 with
// unsigned 64-bit products and long double arithmetic (:1281-1284).
long double g_from_ratios(const uint64_t g[6]) {
    long double G = 66743;
    G *= g[0] * g[2] * g[2] * g[2] * g[4] * g[4];
    G /= g[1] * g[3] * g[3] * g[3] * g[5] * g[5];
    G /= 1000000000000000ULL;
    return G;
}

void set_err(nb_system *s, const std::string &m) {
    if (s) s->err = m;
    else set_static_error(m);
}

template <typename F>
int nsys_guarded(nb_system *s, F &&f) {
    try {
        return f();
    } catch (const NsysError &e) {
        set_err(s, e.msg);
        return e.code;
    } catch (const ApiError &e) {
        set_err(s, e.what);
        return e.code;
    } catch (const DeviceError &e) {
        set_err(s, e.what);
        return NB_ERR_DEVICE;
    } catch (const std::bad_alloc &) {
        set_err(s, "host allocation failed");
        return NB_ERR_CAPACITY;
    }
}

void fill_info(const Parsed &P, nb_nsys_info *info) {
    info->d_types = P.d_types;
    info->time_size = P.time_size;
    info->m_size = P.m_size;
    info->r_size = P.r_size;
    info->t_size = P.t_size;
    info->rest_size = P.rest_size;
    std::memcpy(info->g_vals, P.g, sizeof P.g);
    info->num_bodies = P.n;
    info->time_elapsed = (double)P.time;
    info->dt = (double)P.dt;
}

}  // namespace

extern "C" {

int nb_nsys_read(const char *path, int check_nan_inf, nb_nsys_info *info, int64_t cap, double *m, double *r,
                 double *pos, double *vel, double *rest_c) {
    return nsys_guarded(nullptr, [&]() {
        if (!path || !info) throw ApiError(NB_ERR_INVALID_ARGUMENT, "nb_nsys_read: NULL argument");
        Parsed P;
        read_nsys(path, check_nan_inf != 0, false, P);
        fill_info(P, info);
        if (cap < (int64_t)P.n) throw ApiError(NB_ERR_CAPACITY, "nb_nsys_read: capacity smaller than the body count");
        if (P.n == 0) return NB_OK;
        if (!m || !pos || !vel) throw ApiError(NB_ERR_INVALID_ARGUMENT, "nb_nsys_read: NULL body array");
        read_nsys(path, check_nan_inf != 0, true, P);
        const size_t n = (size_t)P.n;
        std::memcpy(m, P.m.data(), n * 8);
        if (r) std::memcpy(r, P.r.data(), n * 8);
        std::memcpy(pos, P.pos.data(), 3 * n * 8);
        std::memcpy(vel, P.vel.data(), 3 * n * 8);
        if (rest_c) std::memcpy(rest_c, P.rest.data(), n * 8);
        return NB_OK;
    });
}

int nb_nsys_write(const char *path, const nb_nsys_info *info, int64_t n, const double *m, const double *r,
                  const double *pos, const double *vel, const double *rest_c, int value_size, int64_t *bytes) {
    return nsys_guarded(nullptr, [&]() {
        if (!path || !info) throw ApiError(NB_ERR_INVALID_ARGUMENT, "nb_nsys_write: NULL argument");
        const int64_t b = write_nsys(path, (long double)info->time_elapsed, (long double)info->dt, info->g_vals, n, m,
                                     r, pos, vel, rest_c, value_size);
        if (bytes) *bytes = b;
        return NB_OK;
    });
}

int nb_load_nsys(nb_system *s, const char *path, int check_nan_inf, int allow_overlapping) {
    if (!s) return NB_ERR_INVALID_ARGUMENT;
    return nsys_guarded(s, [&]() {
        if (!path) throw ApiError(NB_ERR_INVALID_ARGUMENT, "nb_load_nsys: NULL path");
        Parsed P;
        read_nsys(path, check_nan_inf != 0, true, P);
        const int rc = nb_set_bodies(s, (int64_t)P.n, P.m.data(), P.r.data(), P.pos.data(), P.vel.data(),
                                     P.rest.data(), allow_overlapping ? 0 : 1);
        if (rc != NB_OK) return rc;  // s->err holds the message
        s->dt = (double)P.dt;
        s->half_dt = s->dt / 2;
        std::memcpy(s->G_vals, P.g, sizeof P.g);
        s->G = (double)g_from_ratios(P.g);
        s->time_elapsed = P.time;
        s->acc_valid = false;
        return NB_OK;
    });
}

int nb_save_nsys(const nb_system *cs, const char *path, int value_size, int64_t *bytes) {
    nb_system *s = const_cast<nb_system *>(cs);
    if (!s) return NB_ERR_INVALID_ARGUMENT;
    return nsys_guarded(s, [&]() {
        if (!path) throw ApiError(NB_ERR_INVALID_ARGUMENT, "nb_save_nsys: NULL path");
        const size_t n = (size_t)s->n;
        std::vector<double> m(n), r(n), pos(3 * n), vel(3 * n), rest(n);
        if (n) {
            const int rc = nb_get_state(s, (int64_t)n, m.data(), r.data(), pos.data(), vel.data(), nullptr,
                                        rest.data());
            if (rc != NB_OK) return rc;
        }
        const int64_t b = write_nsys(path, s->time_elapsed, (long double)s->dt, s->G_vals, (int64_t)n, m.data(),
                                     r.data(), pos.data(), vel.data(), rest.data(), value_size);
        if (bytes) *bytes = b;
        return NB_OK;
    });
}

int nb_time_elapsed(const nb_system *s, double *t) {
    if (!s || !t) return NB_ERR_INVALID_ARGUMENT;
    *t = (double)s->time_elapsed;
    return NB_OK;
}

}  // extern "C"
