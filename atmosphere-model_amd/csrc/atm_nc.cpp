// atm_nc.cpp -- netCDF classic-format files (CDF-1, CDF-2, CDF-5) on the
// host: the model's initial, output and restart files without the netCDF
// library (the reference's I/O goes through netCDF-Fortran,
// src/mod_netcdf.f90:1-632).
//
// File layout, per the published netCDF classic format specification:
//   header = magic numrecs dim_list gatt_list var_list
//   data   = the non-record variables in definition order, each from its
//            `begin`, then the records: record r of record variable v
//            starts at begin_v + r * recsize.
// Header integers and all values are big-endian.  CDF-1 stores offsets in
// 32 bits and CDF-2 in 64; CDF-5 also widens every count and size to 64.
//
// The file is the only state: a header is parsed on open, kept in memory,
// and rewritten at enddef; values move with pread/pwrite.  Leaving define
// mode after adding definitions moves every value already written to its
// new offset (the header grew), as the netCDF library does.
#include "../../include/atm_nc.h"

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

namespace {

constexpr uint32_t TAG_DIM = 0x0A, TAG_VAR = 0x0B, TAG_ATT = 0x0C;
constexpr size_t CHUNK = size_t(1) << 20;  // elements per pread/pwrite of a run

int type_size(int t) {
    switch (t) {
    case ATM_NC_BYTE: case ATM_NC_CHAR: case ATM_NC_UBYTE: return 1;
    case ATM_NC_SHORT: case ATM_NC_USHORT: return 2;
    case ATM_NC_INT: case ATM_NC_FLOAT: case ATM_NC_UINT: return 4;
    case ATM_NC_DOUBLE: case ATM_NC_INT64: case ATM_NC_UINT64: return 8;
    }
    return 0;
}
bool type_ok(int t, int version) { return t >= ATM_NC_BYTE && t <= (version == 5 ? ATM_NC_UINT64 : ATM_NC_DOUBLE); }
uint64_t pad4(uint64_t n) { return (n + 3) & ~uint64_t(3); }

void put_be(uint8_t *p, uint64_t v, int n) {
    for (int i = n - 1; i >= 0; i--) {
        p[i] = uint8_t(v);
        v >>= 8;
    }
}
uint64_t get_be(const uint8_t *p, int n) {
    uint64_t v = 0;
    for (int i = 0; i < n; i++) v = (v << 8) | p[i];
    return v;
}

double decode(const uint8_t *p, int t) {
    switch (t) {
    case ATM_NC_BYTE: return double(int8_t(p[0]));
    case ATM_NC_CHAR: case ATM_NC_UBYTE: return double(p[0]);
    case ATM_NC_SHORT: return double(int16_t(get_be(p, 2)));
    case ATM_NC_USHORT: return double(uint16_t(get_be(p, 2)));
    case ATM_NC_INT: return double(int32_t(get_be(p, 4)));
    case ATM_NC_UINT: return double(uint32_t(get_be(p, 4)));
    case ATM_NC_FLOAT: {
        const uint32_t u = uint32_t(get_be(p, 4));
        float f;
        std::memcpy(&f, &u, 4);
        return double(f);
    }
    case ATM_NC_DOUBLE: {
        const uint64_t u = get_be(p, 8);
        double d;
        std::memcpy(&d, &u, 8);
        return d;
    }
    case ATM_NC_INT64: return double(int64_t(get_be(p, 8)));
    case ATM_NC_UINT64: return double(get_be(p, 8));
    }
    return 0.;
}

// integer conversion: truncation toward zero, as a C cast; out of range
// (or NaN) stores the clamped value and reports false
template <class T>
bool to_int(double v, T &out) {
    const double lo = double(std::numeric_limits<T>::min()), hi = double(std::numeric_limits<T>::max());
    if (!(v >= lo)) {
        out = std::numeric_limits<T>::min();
        return false;
    }
    if (!(v <= hi) || (sizeof(T) == 8 && v >= hi)) {  // 2^63, 2^64 are not representable
        out = std::numeric_limits<T>::max();
        return false;
    }
    out = T(v);
    return true;
}

bool encode(double v, uint8_t *p, int t) {
    bool ok = true;
    switch (t) {
    case ATM_NC_BYTE: { int8_t x; ok = to_int(v, x); p[0] = uint8_t(x); break; }
    case ATM_NC_CHAR: case ATM_NC_UBYTE: { uint8_t x; ok = to_int(v, x); p[0] = x; break; }
    case ATM_NC_SHORT: { int16_t x; ok = to_int(v, x); put_be(p, uint16_t(x), 2); break; }
    case ATM_NC_USHORT: { uint16_t x; ok = to_int(v, x); put_be(p, x, 2); break; }
    case ATM_NC_INT: { int32_t x; ok = to_int(v, x); put_be(p, uint32_t(x), 4); break; }
    case ATM_NC_UINT: { uint32_t x; ok = to_int(v, x); put_be(p, x, 4); break; }
    case ATM_NC_INT64: { int64_t x; ok = to_int(v, x); put_be(p, uint64_t(x), 8); break; }
    case ATM_NC_UINT64: { uint64_t x; ok = to_int(v, x); put_be(p, x, 8); break; }
    case ATM_NC_FLOAT: {
        float f;
        if (std::isfinite(v) && std::fabs(v) > double(FLT_MAX)) {
            ok = false;
            f = v > 0 ? INFINITY : -INFINITY;
        } else {
            f = float(v);  // round to nearest even
        }
        uint32_t u;
        std::memcpy(&u, &f, 4);
        put_be(p, u, 4);
        break;
    }
    case ATM_NC_DOUBLE: {
        uint64_t u;
        std::memcpy(&u, &v, 8);
        put_be(p, u, 8);
        break;
    }
    }
    return ok;
}

// n values at once, the type dispatched once: big-endian bytes <-> double
void decode_n(const uint8_t *p, size_t n, int t, double *out) {
    switch (t) {
    case ATM_NC_DOUBLE:
        for (size_t i = 0; i < n; i++) {
            uint64_t u;
            std::memcpy(&u, p + 8 * i, 8);
            u = __builtin_bswap64(u);
            std::memcpy(out + i, &u, 8);
        }
        return;
    case ATM_NC_FLOAT:
        for (size_t i = 0; i < n; i++) {
            uint32_t u;
            std::memcpy(&u, p + 4 * i, 4);
            u = __builtin_bswap32(u);
            float f;
            std::memcpy(&f, &u, 4);
            out[i] = double(f);
        }
        return;
    case ATM_NC_INT:
        for (size_t i = 0; i < n; i++) {
            uint32_t u;
            std::memcpy(&u, p + 4 * i, 4);
            out[i] = double(int32_t(__builtin_bswap32(u)));
        }
        return;
    }
    const int es = type_size(t);
    for (size_t i = 0; i < n; i++) out[i] = decode(p + i * size_t(es), t);
}

bool encode_n(const double *in, size_t n, uint8_t *p, int t) {
    bool ok = true;
    switch (t) {
    case ATM_NC_DOUBLE:
        for (size_t i = 0; i < n; i++) {
            uint64_t u;
            std::memcpy(&u, in + i, 8);
            u = __builtin_bswap64(u);
            std::memcpy(p + 8 * i, &u, 8);
        }
        return true;
    case ATM_NC_FLOAT:
        for (size_t i = 0; i < n; i++) {
            const double v = in[i];
            if (std::isfinite(v) && std::fabs(v) > double(FLT_MAX)) {
                ok &= encode(v, p + 4 * i, t);
                continue;
            }
            const float f = float(v);
            uint32_t u;
            std::memcpy(&u, &f, 4);
            u = __builtin_bswap32(u);
            std::memcpy(p + 4 * i, &u, 4);
        }
        return ok;
    }
    const int es = type_size(t);
    for (size_t i = 0; i < n; i++) ok &= encode(in[i], p + i * size_t(es), t);
    return ok;
}

// the netCDF default fill values, so other readers see unwritten values as missing
void default_fill(int t, uint8_t *p) {
    switch (t) {
    case ATM_NC_BYTE: p[0] = uint8_t(int8_t(-127)); break;
    case ATM_NC_CHAR: p[0] = 0; break;
    case ATM_NC_UBYTE: p[0] = 255; break;
    case ATM_NC_SHORT: put_be(p, uint16_t(int16_t(-32767)), 2); break;
    case ATM_NC_USHORT: put_be(p, 65535u, 2); break;
    case ATM_NC_INT: put_be(p, uint32_t(int32_t(-2147483647)), 4); break;
    case ATM_NC_UINT: put_be(p, 4294967295u, 4); break;
    case ATM_NC_INT64: put_be(p, uint64_t(int64_t(-9223372036854775806LL)), 8); break;
    case ATM_NC_UINT64: put_be(p, 18446744073709551614ULL, 8); break;
    case ATM_NC_FLOAT: encode(double(9.9692099683868690e+36f), p, t); break;
    case ATM_NC_DOUBLE: encode(9.9692099683868690e+36, p, t); break;
    }
}

struct Att {
    std::string name;
    int type = 0;
    size_t n = 0;
    std::vector<uint8_t> ext;  // n values in the external representation
};
struct Dim {
    std::string name;
    uint64_t len = 0;  // 0: the record dimension
};
struct Var {
    std::string name;
    int type = 0;
    std::vector<int> dims;
    std::vector<Att> atts;
    bool rec = false;
    uint64_t nel = 1;    // elements in one record (record variable) or in all
    uint64_t vsize = 0;  // nel * type size, padded to 4 bytes
    uint64_t begin = 0;
};
struct File {
    int fd = -1;
    bool writable = false, define = false, laid_out = false;
    int version = 1;
    uint64_t numrecs = 0, recsize = 0;
    int unlim = -1;
    std::vector<Dim> dims;
    std::vector<Att> gatts;
    std::vector<Var> vars;
    std::vector<Var> committed;  // the layout on disk when define mode was entered
    uint64_t committed_recsize = 0;
    int nw() const { return version == 5 ? 8 : 4; }  // width of counts and sizes
    int ow() const { return version == 1 ? 4 : 8; }  // width of offsets
};

std::mutex g_mu;
std::vector<std::unique_ptr<File>> g_files;

File *lookup(int id) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (id < 0 || size_t(id) >= g_files.size()) return nullptr;
    return g_files[size_t(id)].get();
}
int install(std::unique_ptr<File> f) {
    std::lock_guard<std::mutex> lk(g_mu);
    for (size_t i = 0; i < g_files.size(); i++)
        if (!g_files[i]) {
            g_files[i] = std::move(f);
            return int(i);
        }
    g_files.push_back(std::move(f));
    return int(g_files.size() - 1);
}
void release(int id) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_files[size_t(id)].reset();
}

bool read_all(int fd, void *buf, size_t n, uint64_t off, size_t *got) {
    size_t done = 0;
    while (done < n) {
        const ssize_t r = pread(fd, static_cast<uint8_t *>(buf) + done, n - done, off_t(off + done));
        if (r < 0) {
            if (errno == EINTR) continue;
            return false;
        }
        if (r == 0) break;
        done += size_t(r);
    }
    *got = done;
    return true;
}
bool write_all(int fd, const void *buf, size_t n, uint64_t off) {
    size_t done = 0;
    while (done < n) {
        const ssize_t r = pwrite(fd, static_cast<const uint8_t *>(buf) + done, n - done, off_t(off + done));
        if (r < 0) {
            if (errno == EINTR) continue;
            return false;
        }
        done += size_t(r);
    }
    return true;
}

bool valid_name(const char *s) {
    if (!s || !*s) return false;
    const size_t n = std::strlen(s);
    if (n > ATM_NC_MAX_NAME) return false;
    for (size_t i = 0; i < n; i++) {
        const unsigned char c = static_cast<unsigned char>(s[i]);
        if (c < 0x20 || c == 0x7f || c == '/') return false;
    }
    return true;
}

// ---- shapes and layout ----------------------------------------------------

void compute_shapes(File &f) {
    f.recsize = 0;
    int nrec = 0;
    const Var *last_rec = nullptr;
    for (Var &v : f.vars) {
        v.rec = !v.dims.empty() && v.dims[0] == f.unlim;
        v.nel = 1;
        for (size_t d = v.rec ? 1 : 0; d < v.dims.size(); d++) v.nel *= f.dims[size_t(v.dims[d])].len;
        v.vsize = pad4(v.nel * uint64_t(type_size(v.type)));
        if (v.rec) {
            f.recsize += v.vsize;
            nrec++;
            last_rec = &v;
        }
    }
    // a lone record variable's records are not padded
    if (nrec == 1) f.recsize = last_rec->nel * uint64_t(type_size(last_rec->type));
}

struct Writer {
    std::vector<uint8_t> b;
    void u(uint64_t v, int w) {
        const size_t o = b.size();
        b.resize(o + size_t(w));
        put_be(&b[o], v, w);
    }
    void raw(const uint8_t *p, size_t n) {
        b.insert(b.end(), p, p + n);
        b.resize(size_t(pad4(b.size())), 0);
    }
    void name(const std::string &s, int nw) {
        u(s.size(), nw);
        raw(reinterpret_cast<const uint8_t *>(s.data()), s.size());
    }
};

void put_atts(Writer &w, const std::vector<Att> &atts, int nw) {
    if (atts.empty()) {
        w.u(0, 4);
        w.u(0, nw);
        return;
    }
    w.u(TAG_ATT, 4);
    w.u(atts.size(), nw);
    for (const Att &a : atts) {
        w.name(a.name, nw);
        w.u(uint64_t(a.type), 4);
        w.u(a.n, nw);
        w.raw(a.ext.data(), a.ext.size());
    }
}

std::vector<uint8_t> serialize(const File &f) {
    Writer w;
    const int nw = f.nw(), ow = f.ow();
    const uint8_t magic[4] = {'C', 'D', 'F', uint8_t(f.version)};
    w.b.assign(magic, magic + 4);
    w.u(f.numrecs, nw);
    if (f.dims.empty()) {
        w.u(0, 4);
        w.u(0, nw);
    } else {
        w.u(TAG_DIM, 4);
        w.u(f.dims.size(), nw);
        for (const Dim &d : f.dims) {
            w.name(d.name, nw);
            w.u(d.len, nw);
        }
    }
    put_atts(w, f.gatts, nw);
    if (f.vars.empty()) {
        w.u(0, 4);
        w.u(0, nw);
    } else {
        w.u(TAG_VAR, 4);
        w.u(f.vars.size(), nw);
        const uint64_t vmax = nw == 4 ? 0xFFFFFFFFull : ~uint64_t(0);
        for (const Var &v : f.vars) {
            w.name(v.name, nw);
            w.u(v.dims.size(), nw);
            for (int d : v.dims) w.u(uint64_t(d), nw);
            put_atts(w, v.atts, nw);
            w.u(uint64_t(v.type), 4);
            w.u(v.vsize > vmax ? vmax : v.vsize, nw);  // sizes beyond 32 bits are stored saturated
            w.u(v.begin, ow);
        }
    }
    return w.b;
}

// begins of the variables after a header of the current definitions
int layout(File &f) {
    compute_shapes(f);
    uint64_t off = serialize(f).size();  // the header's size does not depend on the begins
    for (Var &v : f.vars)
        if (!v.rec) {
            v.begin = off;
            off += v.vsize;
        }
    for (Var &v : f.vars)
        if (v.rec) {
            v.begin = off;
            off += v.vsize;
        }
    const uint64_t omax = f.version == 1 ? 0x7FFFFFFFull : 0x7FFFFFFFFFFFFFFFull;
    for (const Var &v : f.vars)
        if (v.begin > omax) return ATM_NC_EVARSIZE;
    if (f.version != 5) {  // 32-bit sizes: only the last variable may exceed 4 GiB - 4
        for (size_t i = 0; i + 1 < f.vars.size(); i++)
            if (f.vars[i].vsize > 0xFFFFFFFCull) return ATM_NC_EVARSIZE;
    }
    return ATM_NC_NOERR;
}

std::vector<uint8_t> fill_pattern(const Var &v) {
    std::vector<uint8_t> one(static_cast<size_t>(type_size(v.type)));
    default_fill(v.type, one.data());
    for (const Att &a : v.atts)
        if (a.name == "_FillValue" && a.type == v.type && a.n == 1) one = a.ext;
    return one;
}

bool fill_bytes(File &f, const Var &v, uint64_t off, uint64_t nbytes) {
    const std::vector<uint8_t> one = fill_pattern(v);
    const size_t es = one.size();
    std::vector<uint8_t> buf(size_t(std::min<uint64_t>(nbytes, CHUNK * es)));
    for (size_t i = 0; i < buf.size(); i++) buf[i] = one[i % es];
    for (uint64_t done = 0; done < nbytes;) {
        const size_t n = size_t(std::min<uint64_t>(buf.size(), nbytes - done));
        if (!write_all(f.fd, buf.data(), n, off + done)) return false;
        done += n;
    }
    return true;
}

// bytes of one record of a record variable: a lone record variable's
// records are not padded
uint64_t rec_bytes(const Var &v, uint64_t recsize) { return std::min(v.vsize, recsize); }

int write_numrecs(File &f) {
    uint8_t b[8];
    put_be(b, f.numrecs, f.nw());
    return write_all(f.fd, b, size_t(f.nw()), 4) ? ATM_NC_NOERR : ATM_NC_EIO;
}

// records numrecs .. n-1 of every record variable, filled
int add_records(File &f, uint64_t n) {
    for (uint64_t r = f.numrecs; r < n; r++)
        for (const Var &v : f.vars)
            if (v.rec && !fill_bytes(f, v, v.begin + r * f.recsize, rec_bytes(v, f.recsize))) return ATM_NC_EIO;
    f.numrecs = n;
    return write_numrecs(f);
}

// ---- header parsing -------------------------------------------------------

struct Reader {
    const uint8_t *p;
    size_t n, i = 0;
    bool short_ = false;
    uint64_t u(int w) {
        if (i + size_t(w) > n) {
            short_ = true;
            i = n;
            return 0;
        }
        const uint64_t v = get_be(p + i, w);
        i += size_t(w);
        return v;
    }
    bool bytes(size_t len, std::vector<uint8_t> &out) {
        if (len > n - i || pad4(len) > n - i) {
            short_ = true;
            i = n;
            return false;
        }
        out.assign(p + i, p + i + len);
        i += size_t(pad4(len));
        return true;
    }
    bool name(std::string &s, int nw) {
        const uint64_t len = u(nw);
        if (short_ || len > ATM_NC_MAX_NAME) return false;
        std::vector<uint8_t> b;
        if (!bytes(size_t(len), b)) return false;
        s.assign(b.begin(), b.end());
        return true;
    }
};

// 1: parsed, 0: need more bytes, <0: status
int parse_atts(Reader &r, std::vector<Att> &atts, int nw, int version) {
    const uint64_t tag = r.u(4), n = r.u(nw);
    if (r.short_) return 0;
    if (tag == 0 && n == 0) return 1;
    if (tag != TAG_ATT || n > (1u << 20)) return ATM_NC_ENOTNC;
    for (uint64_t k = 0; k < n; k++) {
        Att a;
        if (!r.name(a.name, nw)) return r.short_ ? 0 : ATM_NC_ENOTNC;
        a.type = int(r.u(4));
        a.n = size_t(r.u(nw));
        if (r.short_) return 0;
        if (!type_ok(a.type, version)) return ATM_NC_ENOTNC;
        if (a.n > r.n) {  // more values than bytes read so far
            r.short_ = true;
            return 0;
        }
        if (!r.bytes(a.n * size_t(type_size(a.type)), a.ext)) return 0;
        atts.push_back(std::move(a));
    }
    return 1;
}

int parse_header(File &f, const uint8_t *p, size_t n, uint64_t file_size) {
    if (n < 4) return 0;
    static const uint8_t hdf5[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
    if (n >= 8 && std::memcmp(p, hdf5, 8) == 0) return ATM_NC_ENOTNC;
    if (p[0] != 'C' || p[1] != 'D' || p[2] != 'F' || (p[3] != 1 && p[3] != 2 && p[3] != 5)) return ATM_NC_ENOTNC;
    f.version = p[3];
    f.dims.clear();
    f.gatts.clear();
    f.vars.clear();
    f.unlim = -1;
    Reader r{p, n, 4};
    const int nw = f.nw(), ow = f.ow();
    const uint64_t numrecs = r.u(nw);
    const bool streaming = numrecs == (nw == 4 ? 0xFFFFFFFFull : ~uint64_t(0));
    uint64_t tag = r.u(4), cnt = r.u(nw);
    if (r.short_) return 0;
    if (!(tag == 0 && cnt == 0)) {
        if (tag != TAG_DIM || cnt > (1u << 20)) return ATM_NC_ENOTNC;
        for (uint64_t k = 0; k < cnt; k++) {
            Dim d;
            if (!r.name(d.name, nw)) return r.short_ ? 0 : ATM_NC_ENOTNC;
            d.len = r.u(nw);
            if (r.short_) return 0;
            if (d.len == 0) {
                if (f.unlim >= 0) return ATM_NC_ENOTNC;
                f.unlim = int(k);
            }
            f.dims.push_back(d);
        }
    }
    int rc = parse_atts(r, f.gatts, nw, f.version);
    if (rc <= 0) return rc;
    tag = r.u(4);
    cnt = r.u(nw);
    if (r.short_) return 0;
    if (!(tag == 0 && cnt == 0)) {
        if (tag != TAG_VAR || cnt > (1u << 20)) return ATM_NC_ENOTNC;
        for (uint64_t k = 0; k < cnt; k++) {
            Var v;
            if (!r.name(v.name, nw)) return r.short_ ? 0 : ATM_NC_ENOTNC;
            const uint64_t nd = r.u(nw);
            if (r.short_) return 0;
            if (nd > 1024) return ATM_NC_ENOTNC;
            for (uint64_t d = 0; d < nd; d++) {
                const uint64_t id = r.u(nw);
                if (r.short_) return 0;
                if (id >= f.dims.size()) return ATM_NC_ENOTNC;
                if (int(id) == f.unlim && d != 0) return ATM_NC_ENOTNC;  // record dimension first
                v.dims.push_back(int(id));
            }
            rc = parse_atts(r, v.atts, nw, f.version);
            if (rc <= 0) return rc;
            v.type = int(r.u(4));
            r.u(nw);  // vsize: recomputed from the shape
            v.begin = r.u(ow);
            if (r.short_) return 0;
            if (!type_ok(v.type, f.version)) return ATM_NC_ENOTNC;
            f.vars.push_back(std::move(v));
        }
    }
    compute_shapes(f);
    f.numrecs = numrecs;
    if (streaming) {  // records run to the end of the file
        uint64_t first = UINT64_MAX;
        for (const Var &v : f.vars)
            if (v.rec) first = std::min(first, v.begin);
        f.numrecs = (first != UINT64_MAX && f.recsize && file_size > first) ? (file_size - first) / f.recsize : 0;
    }
    return 1;
}

// ---- value transfer ---------------------------------------------------------

int check_var(File *f, int varid, bool write) {
    if (!f) return ATM_NC_EBADID;
    if (f->define) return ATM_NC_EINDEFINE;
    if (varid < 0 || size_t(varid) >= f->vars.size()) return ATM_NC_ENOTVAR;
    if (write && !f->writable) return ATM_NC_EPERM;
    if (f->vars[size_t(varid)].type == ATM_NC_CHAR) return ATM_NC_ECHAR;
    return ATM_NC_NOERR;
}

int transfer(File &f, int varid, const size_t *start, const size_t *count, const ptrdiff_t *imap, double *mem,
             bool write) {
    const Var &v = f.vars[size_t(varid)];
    const int n = int(v.dims.size());
    if (n > 0 && (!start || !count)) return ATM_NC_EINVAL;
    std::vector<uint64_t> len(static_cast<size_t>(n)), cnt(static_cast<size_t>(n)), st(static_cast<size_t>(n));
    uint64_t total = 1;
    for (int d = 0; d < n; d++) {
        const bool recdim = v.rec && d == 0;
        len[size_t(d)] = recdim ? f.numrecs : f.dims[size_t(v.dims[size_t(d)])].len;
        st[size_t(d)] = start[d];
        cnt[size_t(d)] = count[d];
        if (!(recdim && write)) {
            if (st[size_t(d)] > len[size_t(d)]) return ATM_NC_EINVALCOORDS;
            if (cnt[size_t(d)] > len[size_t(d)] - st[size_t(d)]) return ATM_NC_EEDGE;
        }
        total *= cnt[size_t(d)];
    }
    if (total == 0) return ATM_NC_NOERR;
    if (write && v.rec && st[0] + cnt[0] > f.numrecs) {
        const int rc = add_records(f, st[0] + cnt[0]);
        if (rc) return rc;
        len[0] = f.numrecs;
    }
    // memory strides: contiguous C order unless a map is given
    std::vector<ptrdiff_t> mstride(static_cast<size_t>(n));
    for (int d = n - 1; d >= 0; d--)
        mstride[size_t(d)] = imap ? imap[d] : (d == n - 1 ? 1 : mstride[size_t(d) + 1] * ptrdiff_t(cnt[size_t(d) + 1]));
    // file strides (elements) of the dimensions inside one record (or the variable)
    const int first = v.rec ? 1 : 0;
    std::vector<uint64_t> fstride(static_cast<size_t>(n), 0);
    for (int d = n - 1; d >= first; d--)
        fstride[size_t(d)] = d == n - 1 ? 1 : fstride[size_t(d) + 1] * len[size_t(d) + 1];
    // run dimensions k .. n-1: every dimension after k is whole, so a run
    // is contiguous in the file; records are never contiguous
    int k = n;
    if (n > first) {
        k = n - 1;
        while (k > first && st[size_t(k)] == 0 && cnt[size_t(k)] == len[size_t(k)]) k--;
    }
    uint64_t run = 1;
    bool contiguous = true;  // memory strides of the run dimensions are C order
    for (int d = n - 1; d >= k; d--) {
        contiguous &= mstride[size_t(d)] == ptrdiff_t(run);
        run *= cnt[size_t(d)];
    }
    const int es = type_size(v.type);
    std::vector<uint8_t> buf;
    try {
        buf.resize(size_t(std::min<uint64_t>(run, CHUNK)) * size_t(es));
    } catch (const std::bad_alloc &) {
        return ATM_NC_ENOMEM;
    }
    const std::vector<uint8_t> fillv = fill_pattern(v);
    bool range_ok = true;
    std::vector<uint64_t> oi(static_cast<size_t>(k), 0);  // index over the outer dimensions
    std::vector<uint64_t> ri(static_cast<size_t>(n), 0);  // index inside the run
    for (;;) {
        uint64_t off = v.begin;
        ptrdiff_t mbase = 0;
        for (int d = 0; d < k; d++) {
            const uint64_t idx = st[size_t(d)] + oi[size_t(d)];
            if (v.rec && d == 0)
                off += idx * f.recsize;
            else
                off += idx * fstride[size_t(d)] * uint64_t(es);
            mbase += ptrdiff_t(oi[size_t(d)]) * mstride[size_t(d)];
        }
        for (int d = k; d < n; d++) {
            if (!(v.rec && d == 0)) off += st[size_t(d)] * fstride[size_t(d)] * uint64_t(es);
            ri[size_t(d)] = 0;
        }
        ptrdiff_t moff = 0;  // memory offset of the current run element relative to mbase
        for (uint64_t done = 0; done < run;) {
            const size_t m = size_t(std::min<uint64_t>(run - done, CHUNK));
            const uint64_t boff = off + done * uint64_t(es);
            if (!write) {
                size_t got = 0;
                if (!read_all(f.fd, buf.data(), m * size_t(es), boff, &got)) return ATM_NC_EIO;
                for (size_t b = got; b < m * size_t(es); b++) buf[b] = fillv[b % size_t(es)];  // unwritten: fill
            }
            if (contiguous) {  // the run is consecutive in memory as well
                if (write)
                    range_ok &= encode_n(mem + mbase + moff, m, buf.data(), v.type);
                else
                    decode_n(buf.data(), m, v.type, mem + mbase + moff);
                moff += ptrdiff_t(m);
            } else {
                for (size_t e = 0; e < m; e++) {
                    uint8_t *p = buf.data() + e * size_t(es);
                    if (write)
                        range_ok &= encode(mem[mbase + moff], p, v.type);
                    else
                        mem[mbase + moff] = decode(p, v.type);
                    for (int d = n - 1; d >= k; d--) {  // next element of the run
                        moff += mstride[size_t(d)];
                        if (++ri[size_t(d)] < cnt[size_t(d)]) break;
                        moff -= ptrdiff_t(cnt[size_t(d)]) * mstride[size_t(d)];
                        ri[size_t(d)] = 0;
                    }
                }
            }
            if (write && !write_all(f.fd, buf.data(), m * size_t(es), boff)) return ATM_NC_EIO;
            done += m;
        }
        int d = k - 1;
        for (; d >= 0; d--) {
            if (++oi[size_t(d)] < cnt[size_t(d)]) break;
            oi[size_t(d)] = 0;
        }
        if (d < 0) break;
    }
    return range_ok ? ATM_NC_NOERR : ATM_NC_ERANGE;
}

Att *find_att(std::vector<Att> &atts, const char *name) {
    for (Att &a : atts)
        if (a.name == name) return &a;
    return nullptr;
}

std::vector<Att> *att_list(File *f, int varid) {
    if (varid == ATM_NC_GLOBAL) return &f->gatts;
    if (varid < 0 || size_t(varid) >= f->vars.size()) return nullptr;
    return &f->vars[size_t(varid)].atts;
}

void copy_name(char *dst, const std::string &s) {
    if (!dst) return;
    std::memcpy(dst, s.data(), s.size());
    dst[s.size()] = '\0';
}

}  // namespace

extern "C" {

int atm_nc_create(const char *path, int cmode, int *ncid) {
    if (!path || !ncid) return ATM_NC_EINVAL;
    const int flags = O_RDWR | O_CREAT | ((cmode & ATM_NC_NOCLOBBER) ? O_EXCL : O_TRUNC);
    const int fd = open(path, flags, 0666);
    if (fd < 0) return errno == EEXIST ? ATM_NC_EEXIST : errno;
    auto f = std::make_unique<File>();
    f->fd = fd;
    f->writable = true;
    f->define = true;
    f->version = (cmode & ATM_NC_64BIT_DATA) ? 5 : (cmode & ATM_NC_64BIT_OFFSET) ? 2 : 1;
    *ncid = install(std::move(f));
    return ATM_NC_NOERR;
}

int atm_nc_open(const char *path, int mode, int *ncid) {
    if (!path || !ncid) return ATM_NC_EINVAL;
    const bool w = (mode & ATM_NC_WRITE) != 0;
    const int fd = open(path, w ? O_RDWR : O_RDONLY);
    if (fd < 0) return errno;
    struct stat sb;
    if (fstat(fd, &sb) != 0) {
        const int e = errno;
        close(fd);
        return e;
    }
    const uint64_t fsize = uint64_t(sb.st_size);
    auto f = std::make_unique<File>();
    f->fd = fd;
    f->writable = w;
    std::vector<uint8_t> head;
    for (size_t want = 8192;; want *= 4) {
        const size_t n = size_t(std::min<uint64_t>(want, fsize));
        head.resize(n);
        size_t got = 0;
        if (!read_all(fd, head.data(), n, 0, &got)) {
            close(fd);
            return ATM_NC_EIO;
        }
        const int rc = parse_header(*f, head.data(), got, fsize);
        if (rc == 1) break;
        if (rc < 0 || n >= fsize) {
            close(fd);
            return rc < 0 ? rc : ATM_NC_ENOTNC;  // header runs past the end of the file
        }
    }
    f->laid_out = true;
    *ncid = install(std::move(f));
    return ATM_NC_NOERR;
}

int atm_nc_redef(int ncid) {
    File *f = lookup(ncid);
    if (!f) return ATM_NC_EBADID;
    if (!f->writable) return ATM_NC_EPERM;
    if (f->define) return ATM_NC_EINDEFINE;
    f->committed = f->vars;
    f->committed_recsize = f->recsize;
    f->define = true;
    return ATM_NC_NOERR;
}

int atm_nc_enddef(int ncid) {
    File *f = lookup(ncid);
    if (!f) return ATM_NC_EBADID;
    if (!f->define) return ATM_NC_ENOTINDEFINE;
    // values written under the old layout, before the header grows
    std::vector<std::vector<uint8_t>> saved;
    if (f->laid_out) {
        saved.resize(f->committed.size());
        for (size_t i = 0; i < f->committed.size(); i++) {
            const Var &v = f->committed[i];
            const uint64_t nrec = v.rec ? f->numrecs : 1;
            const uint64_t nb = v.rec ? rec_bytes(v, f->committed_recsize) : v.vsize;
            try {
                saved[i].resize(size_t(nb * nrec));
            } catch (const std::bad_alloc &) {
                return ATM_NC_ENOMEM;
            }
            for (uint64_t r = 0; r < nrec; r++) {
                size_t got = 0;
                uint8_t *dst = saved[i].data() + r * nb;
                if (!read_all(f->fd, dst, size_t(nb), v.begin + r * f->committed_recsize, &got))
                    return ATM_NC_EIO;
                if (got < nb) {
                    const std::vector<uint8_t> one = fill_pattern(v);
                    for (size_t b = got; b < nb; b++) dst[b] = one[b % one.size()];
                }
            }
        }
    }
    const int rc = layout(*f);
    if (rc) return rc;
    const std::vector<uint8_t> head = serialize(*f);
    if (!write_all(f->fd, head.data(), head.size(), 0)) return ATM_NC_EIO;
    for (size_t i = 0; i < f->vars.size(); i++) {
        const Var &v = f->vars[i];
        const uint64_t nrec = v.rec ? f->numrecs : 1;
        const uint64_t step = v.rec ? f->recsize : 0;
        // a variable's old and new records have the same size: only the
        // lone-record rule could differ, and it depends on the new layout
        const uint64_t nb = v.rec ? rec_bytes(v, f->recsize) : v.vsize;
        const uint64_t nb_old = i < saved.size() && v.rec ? rec_bytes(f->committed[i], f->committed_recsize) : nb;
        for (uint64_t r = 0; r < nrec; r++) {
            const uint64_t off = v.begin + r * step;
            bool ok;
            if (i < saved.size()) {
                const uint64_t n = std::min(nb, nb_old);
                ok = write_all(f->fd, saved[i].data() + r * nb_old, size_t(n), off);
                if (ok && n < nb) ok = fill_bytes(*f, v, off + n, nb - n);
            } else {
                ok = fill_bytes(*f, v, off, nb);
            }
            if (!ok) return ATM_NC_EIO;
        }
    }
    f->define = false;
    f->laid_out = true;
    f->committed.clear();
    return ATM_NC_NOERR;
}

int atm_nc_sync(int ncid) {
    File *f = lookup(ncid);
    if (!f) return ATM_NC_EBADID;
    if (f->define) return ATM_NC_EINDEFINE;
    if (f->writable) {
        const int rc = write_numrecs(*f);
        if (rc) return rc;
        if (fsync(f->fd) != 0) return ATM_NC_EIO;
    }
    return ATM_NC_NOERR;
}

int atm_nc_close(int ncid) {
    File *f = lookup(ncid);
    if (!f) return ATM_NC_EBADID;
    int rc = ATM_NC_NOERR;
    if (f->define) rc = atm_nc_enddef(ncid);
    if (!rc && f->writable) rc = write_numrecs(*f);
    if (close(f->fd) != 0 && !rc) rc = ATM_NC_EIO;
    release(ncid);
    return rc;
}

int atm_nc_inq(int ncid, int *ndims, int *nvars, int *ngatts, int *unlimdimid) {
    File *f = lookup(ncid);
    if (!f) return ATM_NC_EBADID;
    if (ndims) *ndims = int(f->dims.size());
    if (nvars) *nvars = int(f->vars.size());
    if (ngatts) *ngatts = int(f->gatts.size());
    if (unlimdimid) *unlimdimid = f->unlim;
    return ATM_NC_NOERR;
}

int atm_nc_inq_format(int ncid, int *format) {
    File *f = lookup(ncid);
    if (!f) return ATM_NC_EBADID;
    if (format) *format = f->version;
    return ATM_NC_NOERR;
}

int atm_nc_inq_dim(int ncid, int dimid, char *name, size_t *len) {
    File *f = lookup(ncid);
    if (!f) return ATM_NC_EBADID;
    if (dimid < 0 || size_t(dimid) >= f->dims.size()) return ATM_NC_EBADDIM;
    copy_name(name, f->dims[size_t(dimid)].name);
    if (len) *len = size_t(dimid == f->unlim ? f->numrecs : f->dims[size_t(dimid)].len);
    return ATM_NC_NOERR;
}

int atm_nc_inq_dimid(int ncid, const char *name, int *dimid) {
    File *f = lookup(ncid);
    if (!f) return ATM_NC_EBADID;
    if (!name) return ATM_NC_EINVAL;
    for (size_t i = 0; i < f->dims.size(); i++)
        if (f->dims[i].name == name) {
            if (dimid) *dimid = int(i);
            return ATM_NC_NOERR;
        }
    return ATM_NC_EBADDIM;
}

int atm_nc_inq_var(int ncid, int varid, char *name, int *xtype, int *ndims, int *dimids, int *natts) {
    File *f = lookup(ncid);
    if (!f) return ATM_NC_EBADID;
    if (varid < 0 || size_t(varid) >= f->vars.size()) return ATM_NC_ENOTVAR;
    const Var &v = f->vars[size_t(varid)];
    copy_name(name, v.name);
    if (xtype) *xtype = v.type;
    if (ndims) *ndims = int(v.dims.size());
    if (dimids)
        for (size_t d = 0; d < v.dims.size(); d++) dimids[d] = v.dims[d];
    if (natts) *natts = int(v.atts.size());
    return ATM_NC_NOERR;
}

int atm_nc_inq_varid(int ncid, const char *name, int *varid) {
    File *f = lookup(ncid);
    if (!f) return ATM_NC_EBADID;
    if (!name) return ATM_NC_EINVAL;
    for (size_t i = 0; i < f->vars.size(); i++)
        if (f->vars[i].name == name) {
            if (varid) *varid = int(i);
            return ATM_NC_NOERR;
        }
    return ATM_NC_ENOTVAR;
}

int atm_nc_def_dim(int ncid, const char *name, size_t len, int *dimid) {
    File *f = lookup(ncid);
    if (!f) return ATM_NC_EBADID;
    if (!f->define) return ATM_NC_ENOTINDEFINE;
    if (!valid_name(name)) return ATM_NC_EBADNAME;
    for (const Dim &d : f->dims)
        if (d.name == name) return ATM_NC_ENAMEINUSE;
    if (len == ATM_NC_UNLIMITED && f->unlim >= 0) return ATM_NC_EUNLIMIT;
    if (f->version != 5 && uint64_t(len) > 0xFFFFFFFCull) return ATM_NC_EINVAL;
    f->dims.push_back(Dim{name, uint64_t(len)});
    if (len == ATM_NC_UNLIMITED) f->unlim = int(f->dims.size() - 1);
    if (dimid) *dimid = int(f->dims.size() - 1);
    return ATM_NC_NOERR;
}

int atm_nc_def_var(int ncid, const char *name, int xtype, int ndims, const int *dimids, int *varid) {
    File *f = lookup(ncid);
    if (!f) return ATM_NC_EBADID;
    if (!f->define) return ATM_NC_ENOTINDEFINE;
    if (!valid_name(name)) return ATM_NC_EBADNAME;
    if (!type_ok(xtype, f->version)) return ATM_NC_EBADTYPE;
    if (ndims < 0 || (ndims > 0 && !dimids)) return ATM_NC_EINVAL;
    for (const Var &v : f->vars)
        if (v.name == name) return ATM_NC_ENAMEINUSE;
    Var v;
    v.name = name;
    v.type = xtype;
    for (int d = 0; d < ndims; d++) {
        if (dimids[d] < 0 || size_t(dimids[d]) >= f->dims.size()) return ATM_NC_EBADDIM;
        if (dimids[d] == f->unlim && d != 0) return ATM_NC_EUNLIMPOS;
        v.dims.push_back(dimids[d]);
    }
    f->vars.push_back(std::move(v));
    compute_shapes(*f);
    if (varid) *varid = int(f->vars.size() - 1);
    return ATM_NC_NOERR;
}

int atm_nc_inq_att(int ncid, int varid, const char *name, int *xtype, size_t *len) {
    File *f = lookup(ncid);
    if (!f) return ATM_NC_EBADID;
    std::vector<Att> *l = att_list(f, varid);
    if (!l) return ATM_NC_ENOTVAR;
    const Att *a = name ? find_att(*l, name) : nullptr;
    if (!a) return ATM_NC_ENOTATT;
    if (xtype) *xtype = a->type;
    if (len) *len = a->n;
    return ATM_NC_NOERR;
}

int atm_nc_inq_attname(int ncid, int varid, int attnum, char *name) {
    File *f = lookup(ncid);
    if (!f) return ATM_NC_EBADID;
    std::vector<Att> *l = att_list(f, varid);
    if (!l) return ATM_NC_ENOTVAR;
    if (attnum < 0 || size_t(attnum) >= l->size()) return ATM_NC_ENOTATT;
    copy_name(name, (*l)[size_t(attnum)].name);
    return ATM_NC_NOERR;
}

int atm_nc_get_att_double(int ncid, int varid, const char *name, double *values) {
    File *f = lookup(ncid);
    if (!f) return ATM_NC_EBADID;
    std::vector<Att> *l = att_list(f, varid);
    if (!l) return ATM_NC_ENOTVAR;
    const Att *a = name ? find_att(*l, name) : nullptr;
    if (!a) return ATM_NC_ENOTATT;
    if (a->type == ATM_NC_CHAR) return ATM_NC_ECHAR;
    const int es = type_size(a->type);
    for (size_t i = 0; i < a->n; i++) values[i] = decode(a->ext.data() + i * size_t(es), a->type);
    return ATM_NC_NOERR;
}

int atm_nc_get_att_text(int ncid, int varid, const char *name, char *text) {
    File *f = lookup(ncid);
    if (!f) return ATM_NC_EBADID;
    std::vector<Att> *l = att_list(f, varid);
    if (!l) return ATM_NC_ENOTVAR;
    const Att *a = name ? find_att(*l, name) : nullptr;
    if (!a) return ATM_NC_ENOTATT;
    if (a->type != ATM_NC_CHAR) return ATM_NC_ECHAR;
    if (a->n) std::memcpy(text, a->ext.data(), a->n);
    return ATM_NC_NOERR;
}

static int put_att(int ncid, int varid, const char *name, int xtype, size_t len, const double *dv,
                   const char *tv) {
    File *f = lookup(ncid);
    if (!f) return ATM_NC_EBADID;
    if (!f->define) return ATM_NC_ENOTINDEFINE;
    if (!valid_name(name)) return ATM_NC_EBADNAME;
    if (!type_ok(xtype, f->version)) return ATM_NC_EBADTYPE;
    std::vector<Att> *l = att_list(f, varid);
    if (!l) return ATM_NC_ENOTVAR;
    if (len && !dv && !tv) return ATM_NC_EINVAL;
    Att a;
    a.name = name;
    a.type = xtype;
    a.n = len;
    const int es = type_size(xtype);
    a.ext.resize(len * size_t(es));
    bool ok = true;
    for (size_t i = 0; i < len; i++) {
        if (tv)
            a.ext[i] = uint8_t(tv[i]);
        else
            ok &= encode(dv[i], a.ext.data() + i * size_t(es), xtype);
    }
    Att *old = find_att(*l, name);
    if (old)
        *old = std::move(a);
    else
        l->push_back(std::move(a));
    return ok ? ATM_NC_NOERR : ATM_NC_ERANGE;
}

int atm_nc_put_att_double(int ncid, int varid, const char *name, int xtype, size_t len, const double *values) {
    if (xtype == ATM_NC_CHAR) return ATM_NC_ECHAR;
    return put_att(ncid, varid, name, xtype, len, values, nullptr);
}

int atm_nc_put_att_text(int ncid, int varid, const char *name, size_t len, const char *text) {
    return put_att(ncid, varid, name, ATM_NC_CHAR, len, nullptr, text ? text : "");
}

int atm_nc_get_vara_double(int ncid, int varid, const size_t *start, const size_t *count, double *values) {
    File *f = lookup(ncid);
    int rc = check_var(f, varid, false);
    if (rc) return rc;
    if (!values) return ATM_NC_EINVAL;
    return transfer(*f, varid, start, count, nullptr, values, false);
}

int atm_nc_get_varm_double(int ncid, int varid, const size_t *start, const size_t *count, const ptrdiff_t *imap,
                           double *values) {
    File *f = lookup(ncid);
    int rc = check_var(f, varid, false);
    if (rc) return rc;
    if (!values) return ATM_NC_EINVAL;
    return transfer(*f, varid, start, count, imap, values, false);
}

int atm_nc_put_vara_double(int ncid, int varid, const size_t *start, const size_t *count, const double *values) {
    File *f = lookup(ncid);
    int rc = check_var(f, varid, true);
    if (rc) return rc;
    if (!values) return ATM_NC_EINVAL;
    return transfer(*f, varid, start, count, nullptr, const_cast<double *>(values), true);
}

int atm_nc_put_varm_double(int ncid, int varid, const size_t *start, const size_t *count, const ptrdiff_t *imap,
                           const double *values) {
    File *f = lookup(ncid);
    int rc = check_var(f, varid, true);
    if (rc) return rc;
    if (!values) return ATM_NC_EINVAL;
    return transfer(*f, varid, start, count, imap, const_cast<double *>(values), true);
}

const char *atm_nc_strerror(int status) {
    switch (status) {
    case ATM_NC_NOERR: return "No error";
    case ATM_NC_EBADID: return "Not a valid netCDF file id";
    case ATM_NC_EEXIST: return "File exists and NOCLOBBER was given";
    case ATM_NC_EINVAL: return "Invalid argument";
    case ATM_NC_EPERM: return "Write to a file opened read-only";
    case ATM_NC_ENOTINDEFINE: return "Operation needs define mode";
    case ATM_NC_EINDEFINE: return "Operation not allowed in define mode";
    case ATM_NC_EINVALCOORDS: return "Index exceeds dimension bound";
    case ATM_NC_ENAMEINUSE: return "Name already in use";
    case ATM_NC_ENOTATT: return "Attribute not found";
    case ATM_NC_EBADTYPE: return "Unknown or unsupported data type";
    case ATM_NC_EBADDIM: return "Invalid dimension id or name";
    case ATM_NC_EUNLIMPOS: return "The record dimension must be a variable's first dimension";
    case ATM_NC_ENOTVAR: return "Variable not found";
    case ATM_NC_ENOTNC:
        return "Not a netCDF classic-format file (CDF-1, CDF-2 or CDF-5); netCDF-4/HDF5 files are not read";
    case ATM_NC_EUNLIMIT: return "Only one record dimension is allowed";
    case ATM_NC_ECHAR: return "Numeric transfer on a character variable or attribute";
    case ATM_NC_EEDGE: return "Start + count exceeds dimension bound";
    case ATM_NC_EBADNAME: return "Invalid name";
    case ATM_NC_ERANGE: return "Numeric conversion out of the external type's range";
    case ATM_NC_ENOMEM: return "Out of memory";
    case ATM_NC_EVARSIZE: return "Variable or offset too large for the file format";
    case ATM_NC_ETRUNC: return "File shorter than its header says";
    case ATM_NC_EIO: return "Read or write of the file failed";
    }
    if (status > 0) return std::strerror(status);
    return "Unknown status";
}

}  // extern "C"
