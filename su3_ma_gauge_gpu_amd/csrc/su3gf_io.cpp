// su3gf_io.cpp -- configuration files in the formats of the reference's
// applications (include/su3gf_io.h).  Host code: plain C++ file I/O, no GPU.
//
// Every format maps a file element to a host-array element through the
// site's coordinates (t, x, y, z), the direction mu and the matrix entry
// (i, j, c):
//   * memory side: GpuPattern (GpuPattern.hxx:53-55) or GpuPatternTimeslice
//     (GpuPatternTimeslice.hxx:57-61), with the parity-split site index of
//     SiteCoord.hxx:104-196,289-330;
//   * file side: the standard order of StandardPattern.hxx:53-70 (PLAIN,
//     HEADERONLY, VOGT), the ILDG order of ildg.cpp:63-97 or the QCDSTAG order
//     of qcdstag.cpp:43-63.
// The LIME layer (record headers, padding) follows the LIME format that
// c-lime writes (lime_binary_header.h, lime_writer.c:262-318): a 144-byte
// header -- big-endian magic 0x456789ab, big-endian version, MB (0x80) and
// ME (0x40) bits in byte 6, big-endian 64-bit data length, a NUL-padded
// 128-byte type -- then the data, zero-padded to a multiple of 8 bytes.
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "su3gf_io.h"
#include "su3gf_error.h"

namespace {

gf_status io_fail(gf_status st, const std::string &msg) { return su3gf::set_error(st, msg.c_str()); }

// ------------------------------------------------------------ geometry
struct Shape {
    int64_t N[4];  // Nt, Nx, Ny, Nz
    int64_t V, V3;
};

// host-array index of entry (i, j, c) of link (t, x, y, z; mu)
int64_t mem_index(const Shape &s, int pattern, int64_t t, int64_t x, int64_t y, int64_t z, int mu, int i, int j,
                  int c)
{
    const int64_t comp = c + 2 * (j + 3 * (i + 3 * mu));
    if (pattern == GF_PATTERN_GPU_TIMESLICE) {
        const int64_t n3 = (x * s.N[2] + y) * s.N[3] + z;
        const int64_t i3 = n3 / 2 + (((x + y + z) & 1) ? s.V3 / 2 : 0);
        return 72 * s.V3 * t + i3 + s.V3 * comp;
    }
    const int64_t n = ((t * s.N[1] + x) * s.N[2] + y) * s.N[3] + z;
    const int64_t idx = n / 2 + (((t + x + y + z) & 1) ? s.V / 2 : 0);
    return idx + s.V * comp;
}

// host-array index of real u of the standard (site-major) order
int64_t mem_index_standard(const Shape &s, int pattern, int64_t u)
{
    const int c = (int)(u % 2);
    u /= 2;
    const int j = (int)(u % 3);
    u /= 3;
    const int i = (int)(u % 3);
    u /= 3;
    const int mu = (int)(u % 4);
    int64_t n = u / 4;
    const int64_t z = n % s.N[3];
    n /= s.N[3];
    const int64_t y = n % s.N[2];
    n /= s.N[2];
    const int64_t x = n % s.N[1];
    const int64_t t = n / s.N[1];
    return mem_index(s, pattern, t, x, y, z, mu, i, j, c);
}

// ------------------------------------------------------------ host array
struct HostArray {
    void *p;
    bool f64;
    double get(int64_t k) const { return f64 ? ((const double *)p)[k] : (double)((const float *)p)[k]; }
    void set(int64_t k, double v) const
    {
        if (f64) ((double *)p)[k] = v;
        else ((float *)p)[k] = (float)v;
    }
};

// ------------------------------------------------------------ byte helpers
uint64_t bswap64(uint64_t v) { return __builtin_bswap64(v); }

double be_double(const unsigned char *b)
{
    uint64_t v;
    std::memcpy(&v, b, 8);
    v = bswap64(v);
    double d;
    std::memcpy(&d, &v, 8);
    return d;
}

void put_be_double(unsigned char *b, double d)
{
    uint64_t v;
    std::memcpy(&v, &d, 8);
    v = bswap64(v);
    std::memcpy(b, &v, 8);
}

struct File {
    FILE *fp = nullptr;
    ~File()
    {
        if (fp) std::fclose(fp);
    }
};

gf_status read_all(const char *path, std::vector<unsigned char> &out)
{
    File f;
    f.fp = std::fopen(path, "rb");
    if (!f.fp) return io_fail(GF_ERR_IO, std::string("cannot open ") + path + ": " + std::strerror(errno));
    if (std::fseek(f.fp, 0, SEEK_END) != 0) return io_fail(GF_ERR_IO, std::string("cannot seek in ") + path);
    const long len = std::ftell(f.fp);
    if (len < 0) return io_fail(GF_ERR_IO, std::string("cannot size ") + path);
    std::rewind(f.fp);
    out.resize((size_t)len);
    if (len > 0 && std::fread(out.data(), 1, (size_t)len, f.fp) != (size_t)len)
        return io_fail(GF_ERR_IO, std::string("cannot read ") + path);
    return GF_OK;
}

gf_status write_all(const char *path, const std::vector<unsigned char> &data)
{
    File f;
    f.fp = std::fopen(path, "wb");
    if (!f.fp) return io_fail(GF_ERR_IO, std::string("cannot create ") + path + ": " + std::strerror(errno));
    if (!data.empty() && std::fwrite(data.data(), 1, data.size(), f.fp) != data.size())
        return io_fail(GF_ERR_IO, std::string("cannot write ") + path);
    if (std::fclose(f.fp) != 0) {
        f.fp = nullptr;
        return io_fail(GF_ERR_IO, std::string("cannot close ") + path);
    }
    f.fp = nullptr;
    return GF_OK;
}

// ------------------------------------------------------------ LIME records
constexpr uint32_t kLimeMagic = 0x456789abu;
constexpr size_t kLimeHeader = 144;
const char *const kIldgBinary = "ildg-binary-data";

struct LimeRecord {
    std::string type;
    bool mb = false, me = false;
    std::vector<unsigned char> data;
};

size_t lime_pad(size_t n) { return (8 - n % 8) % 8; }

gf_status parse_lime(const std::vector<unsigned char> &buf, const char *path, std::vector<LimeRecord> &recs)
{
    recs.clear();
    size_t pos = 0;
    while (pos < buf.size()) {
        if (buf.size() - pos < kLimeHeader)
            return io_fail(GF_ERR_IO, std::string(path) + ": truncated LIME record header");
        const unsigned char *h = buf.data() + pos;
        const uint32_t magic = (uint32_t)h[0] << 24 | (uint32_t)h[1] << 16 | (uint32_t)h[2] << 8 | h[3];
        if (magic != kLimeMagic) return io_fail(GF_ERR_IO, std::string(path) + ": not a LIME file (bad magic number)");
        uint64_t len = 0;
        for (int k = 0; k < 8; k++) len = len << 8 | h[8 + k];
        LimeRecord r;
        r.mb = (h[6] & 0x80) != 0;
        r.me = (h[6] & 0x40) != 0;
        const char *type = (const char *)h + 16;
        r.type.assign(type, strnlen(type, 128));
        pos += kLimeHeader;
        if (len > buf.size() - pos) return io_fail(GF_ERR_IO, std::string(path) + ": truncated LIME record " + r.type);
        r.data.assign(buf.begin() + pos, buf.begin() + pos + len);
        pos += len;
        pos += std::min(lime_pad(len), buf.size() - pos);
        recs.push_back(std::move(r));
    }
    return GF_OK;
}

void append_lime(std::vector<unsigned char> &out, const LimeRecord &r)
{
    unsigned char h[kLimeHeader] = {};
    h[0] = kLimeMagic >> 24;
    h[1] = (kLimeMagic >> 16) & 0xff;
    h[2] = (kLimeMagic >> 8) & 0xff;
    h[3] = kLimeMagic & 0xff;
    h[5] = 1;  // version 1, big-endian
    h[6] = (unsigned char)((r.mb ? 0x80 : 0) | (r.me ? 0x40 : 0));
    const uint64_t len = r.data.size();
    for (int k = 0; k < 8; k++) h[8 + k] = (unsigned char)(len >> (56 - 8 * k));
    std::memcpy(h + 16, r.type.data(), std::min<size_t>(r.type.size(), 127));
    out.insert(out.end(), h, h + kLimeHeader);
    out.insert(out.end(), r.data.begin(), r.data.end());
    out.insert(out.end(), lime_pad(r.data.size()), 0);
}

// ------------------------------------------------------------ the handle
}  // namespace

struct gf_linkfile {
    int type, reinterpret, pattern;
    bool have_header = false;
    std::vector<unsigned char> header;     // HEADERONLY: raw bytes; VOGT: the header as read
    std::vector<LimeRecord> records;       // ILDG: the records of the last file (binary data dropped)
};

namespace {

size_t real_bytes(const gf_linkfile *f, const gf_lattice *lat)
{
    if (f->reinterpret == GF_REAL_DOUBLE) return 8;
    if (f->reinterpret == GF_REAL_FLOAT) return 4;
    return lat->dtype == GF_F64 ? 8 : 4;
}

gf_status check_args(const gf_linkfile *f, const char *path, const gf_lattice *lat, const void *U, Shape &s)
{
    if (!f || !path || !lat || !U) return io_fail(GF_ERR_BAD_ARG, "null argument");
    if (lat->dtype != GF_F32 && lat->dtype != GF_F64) return io_fail(GF_ERR_BAD_ARG, "unknown dtype");
    if (lat->layout != GF_LAYOUT_PATTERN18)
        return io_fail(GF_ERR_BAD_ARG, "configuration files hold all 18 reals of a link (GF_LAYOUT_PATTERN18)");
    s.V = 1;
    for (int i = 0; i < 4; i++) {
        if (lat->dims[i] <= 0 || lat->dims[i] % 2) return io_fail(GF_ERR_BAD_SHAPE, "every extent must be even and positive");
        s.N[i] = lat->dims[i];
        s.V *= s.N[i];
    }
    s.V3 = s.V / s.N[0];
    return GF_OK;
}

// ---- standard order (PLAIN, HEADERONLY, VOGT): LinkFile.hxx:119-150, 219-252
void decode_standard(const gf_linkfile *f, const Shape &s, const unsigned char *data, size_t rb, const HostArray &U)
{
    const int64_t n = 72 * s.V;
    for (int64_t u = 0; u < n; u++) {
        double v;
        if (rb == 8) {
            std::memcpy(&v, data + 8 * u, 8);
        } else {
            float w;
            std::memcpy(&w, data + 4 * u, 4);
            v = w;
        }
        U.set(mem_index_standard(s, f->pattern, u), v);
    }
}

void encode_standard(const gf_linkfile *f, const Shape &s, const HostArray &U, size_t rb, unsigned char *data)
{
    const int64_t n = 72 * s.V;
    for (int64_t u = 0; u < n; u++) {
        const double v = U.get(mem_index_standard(s, f->pattern, u));
        if (rb == 8) {
            std::memcpy(data + 8 * u, &v, 8);
        } else {
            const float w = (float)v;
            std::memcpy(data + 4 * u, &w, 4);
        }
    }
}

// ---- VOGT header: FileVogt.hxx:52-102
gf_status vogt_header(const std::vector<unsigned char> &buf, const char *path, const Shape &s, size_t rb,
                      size_t &hlen)
{
    auto sh = [&](size_t k) {
        int16_t v;
        std::memcpy(&v, buf.data() + 2 * k, 2);
        return (int)v;
    };
    if (buf.size() < 4) return io_fail(GF_ERR_IO, std::string(path) + ": no VOGT header");
    const int ndim = sh(0), nc = sh(1);
    if (ndim != 4 || nc != 3)
        return io_fail(GF_ERR_IO, std::string(path) + ": VOGT header says ndim " + std::to_string(ndim) + ", nc " +
                                      std::to_string(nc) + " (expected 4 and 3)");
    hlen = 2 * (size_t)(3 + ndim);
    if (buf.size() < hlen) return io_fail(GF_ERR_IO, std::string(path) + ": truncated VOGT header");
    int64_t vol = 1;
    for (int i = 0; i < ndim; i++) vol *= sh(2 + i);
    if (vol != s.V)
        return io_fail(GF_ERR_BAD_SHAPE, std::string(path) + ": VOGT header volume " + std::to_string(vol) +
                                             " differs from the lattice's " + std::to_string(s.V));
    const int len = sh(2 + ndim);
    if (len != (int)rb)
        return io_fail(GF_ERR_IO, std::string(path) + ": VOGT header has " + std::to_string(len) +
                                      "-byte reals, reading " + std::to_string(rb) + "-byte reals");
    return GF_OK;
}

std::vector<unsigned char> default_vogt_header(const Shape &s, size_t rb)
{
    const int16_t v[7] = {4, 3, (int16_t)s.N[0], (int16_t)s.N[1], (int16_t)s.N[2], (int16_t)s.N[3], (int16_t)rb};
    std::vector<unsigned char> h(sizeof v);
    std::memcpy(h.data(), v, sizeof v);
    return h;
}

// ---- ILDG binary record: ildg.cpp:63-97 (read), 122-146 (write)
int64_t ildg_element(const Shape &s, int64_t t, int64_t x, int64_t y, int64_t z, int m, int i, int j, int c)
{
    return ((((t * s.N[3] + z) * s.N[2] + y) * s.N[1] + x) * 4 + m) * 18 + 6 * i + 2 * j + c;
}

template <typename F> void for_each_link_entry(const Shape &s, F &&fn)
{
    for (int64_t t = 0; t < s.N[0]; t++)
        for (int64_t z = 0; z < s.N[3]; z++)
            for (int64_t y = 0; y < s.N[2]; y++)
                for (int64_t x = 0; x < s.N[1]; x++)
                    for (int m = 0; m < 4; m++)
                        for (int i = 0; i < 3; i++)
                            for (int j = 0; j < 3; j++)
                                for (int c = 0; c < 2; c++) fn(t, x, y, z, m, i, j, c);
}

void decode_ildg(const gf_linkfile *f, const Shape &s, const unsigned char *data, const HostArray &U)
{
    for_each_link_entry(s, [&](int64_t t, int64_t x, int64_t y, int64_t z, int m, int i, int j, int c) {
        const int mu = (m + 1) % 4;  // file order x, y, z, t
        U.set(mem_index(s, f->pattern, t, x, y, z, mu, i, j, c), be_double(data + 8 * ildg_element(s, t, x, y, z, m, i, j, c)));
    });
}

void encode_ildg(const gf_linkfile *f, const Shape &s, const HostArray &U, unsigned char *data)
{
    for_each_link_entry(s, [&](int64_t t, int64_t x, int64_t y, int64_t z, int m, int i, int j, int c) {
        const int mu = (m + 1) % 4;
        put_be_double(data + 8 * ildg_element(s, t, x, y, z, m, i, j, c), U.get(mem_index(s, f->pattern, t, x, y, z, mu, i, j, c)));
    });
}

// ---- QCDSTAG: qcdstag.cpp:43-63 (read), 84-106 (write)
int64_t qcdstag_element(const Shape &s, int64_t t, int64_t x, int64_t y, int64_t z, int mu, int i, int j, int c)
{
    return ((((mu * s.N[0] + t) * s.N[3] + z) * s.N[2] + y) * s.N[1] + x) * 18 + 6 * i + 2 * j + c;
}

// ILDG with no template: the ildg-format record of the ILDG standard
std::string ildg_format_xml(const Shape &s)
{
    return "<?xml version=\"1.0\" encoding=\"UTF-8\"?>\n"
           "<ildgFormat xmlns=\"http://www.lqcd.org/ildg\" "
           "xmlns:xsi=\"http://www.w3.org/2001/XMLSchema-instance\" "
           "xsi:schemaLocation=\"http://www.lqcd.org/ildg http://www.lqcd.org/ildg/ildg-format.xsd\">\n"
           "<version>1.0</version>\n<field>su3gauge</field>\n<precision>64</precision>\n"
           "<lx>" + std::to_string(s.N[1]) + "</lx>\n<ly>" + std::to_string(s.N[2]) + "</ly>\n<lz>" +
           std::to_string(s.N[3]) + "</lz>\n<lt>" + std::to_string(s.N[0]) + "</lt>\n</ildgFormat>";
}

gf_status keep_header(gf_linkfile *f, const char *path, const Shape &s, size_t rb,
                      const std::vector<unsigned char> &buf, size_t &data_off, size_t need)
{
    data_off = 0;
    if (f->type == GF_FILE_HEADERONLY) {
        if (buf.size() < need)
            return io_fail(GF_ERR_IO, std::string(path) + ": shorter (" + std::to_string(buf.size()) +
                                          " bytes) than the configuration (" + std::to_string(need) + ")");
        data_off = buf.size() - need;
        f->header.assign(buf.begin(), buf.begin() + data_off);
        f->have_header = true;
    } else if (f->type == GF_FILE_VOGT) {
        size_t hlen = 0;
        gf_status st = vogt_header(buf, path, s, rb, hlen);
        if (st != GF_OK) return st;
        f->header.assign(buf.begin(), buf.begin() + hlen);
        f->have_header = true;
        data_off = hlen;
    } else if (f->type == GF_FILE_ILDG) {
        std::vector<LimeRecord> recs;
        gf_status st = parse_lime(buf, path, recs);
        if (st != GF_OK) return st;
        f->records.clear();
        for (auto &r : recs) {
            if (r.type == kIldgBinary) r.data.clear();  // replaced on save
            f->records.push_back(std::move(r));
        }
        f->have_header = true;
    }
    return GF_OK;
}

}  // namespace

extern "C" {

gf_status gf_linkfile_create(int filetype, int reinterpret, int pattern, gf_linkfile **out)
{
    if (!out) return io_fail(GF_ERR_BAD_ARG, "out is null");
    *out = nullptr;
    if (filetype < GF_FILE_PLAIN || filetype > GF_FILE_QCDSTAG)
        return io_fail(GF_ERR_BAD_ARG, "unknown file type " + std::to_string(filetype));
    if (reinterpret < GF_REAL_STANDARD || reinterpret > GF_REAL_FLOAT)
        return io_fail(GF_ERR_BAD_ARG, "unknown reinterpretation " + std::to_string(reinterpret));
    if (pattern != GF_PATTERN_GPU && pattern != GF_PATTERN_GPU_TIMESLICE)
        return io_fail(GF_ERR_BAD_ARG, "unknown memory pattern " + std::to_string(pattern));
    gf_linkfile *f = new gf_linkfile();
    f->type = filetype;
    f->reinterpret = reinterpret;
    f->pattern = pattern;
    *out = f;
    return GF_OK;
}

void gf_linkfile_destroy(gf_linkfile *f) { delete f; }

gf_status gf_linkfile_read_header(gf_linkfile *f, const char *path, const gf_lattice *lat)
{
    Shape s;
    int dummy = 0;
    gf_status st = check_args(f, path, lat, &dummy, s);
    if (st != GF_OK) return st;
    std::vector<unsigned char> buf;
    if ((st = read_all(path, buf)) != GF_OK) return st;
    const size_t rb = real_bytes(f, lat);
    size_t off = 0;
    return keep_header(f, path, s, rb, buf, off, (size_t)(72 * s.V) * rb);
}

gf_status gf_linkfile_load(gf_linkfile *f, const char *path, const gf_lattice *lat, void *Uv)
{
    Shape s;
    gf_status st = check_args(f, path, lat, Uv, s);
    if (st != GF_OK) return st;
    const HostArray U{Uv, lat->dtype == GF_F64};
    std::vector<unsigned char> buf;
    if ((st = read_all(path, buf)) != GF_OK) return st;
    const size_t n = (size_t)(72 * s.V);
    switch (f->type) {
    case GF_FILE_PLAIN:
    case GF_FILE_HEADERONLY:
    case GF_FILE_VOGT: {
        const size_t rb = real_bytes(f, lat);
        size_t off = 0;
        if ((st = keep_header(f, path, s, rb, buf, off, n * rb)) != GF_OK) return st;
        if (buf.size() - off < n * rb)
            return io_fail(GF_ERR_IO, std::string(path) + ": holds " + std::to_string(buf.size() - off) +
                                          " bytes of links, the lattice needs " + std::to_string(n * rb));
        decode_standard(f, s, buf.data() + off, rb, U);
        return GF_OK;
    }
    case GF_FILE_ILDG: {
        std::vector<LimeRecord> recs;
        if ((st = parse_lime(buf, path, recs)) != GF_OK) return st;
        const LimeRecord *bin = nullptr;
        for (const auto &r : recs)
            if (r.type == kIldgBinary) bin = &r;
        if (!bin) return io_fail(GF_ERR_IO, std::string(path) + ": no ildg-binary-data record");
        if (bin->data.size() < 8 * n)
            return io_fail(GF_ERR_IO, std::string(path) + ": ildg-binary-data holds " + std::to_string(bin->data.size()) +
                                          " bytes, the lattice needs " + std::to_string(8 * n));
        decode_ildg(f, s, bin->data.data(), U);
        f->records.clear();
        for (auto &r : recs) {
            if (r.type == kIldgBinary) r.data.clear();
            f->records.push_back(std::move(r));
        }
        f->have_header = true;
        return GF_OK;
    }
    case GF_FILE_QCDSTAG: {
        if (buf.size() < 8 * n)
            return io_fail(GF_ERR_IO, std::string(path) + ": holds " + std::to_string(buf.size()) +
                                          " bytes, the lattice needs " + std::to_string(8 * n));
        for_each_link_entry(s, [&](int64_t t, int64_t x, int64_t y, int64_t z, int mu, int i, int j, int c) {
            double v;
            std::memcpy(&v, buf.data() + 8 * qcdstag_element(s, t, x, y, z, mu, i, j, c), 8);
            U.set(mem_index(s, f->pattern, t, x, y, z, mu, i, j, c), v);
        });
        return GF_OK;
    }
    default:
        return io_fail(GF_ERR_BAD_ARG, "unknown file type");
    }
}

gf_status gf_linkfile_save(gf_linkfile *f, const char *path, const gf_lattice *lat, const void *Uv, int sa_steps)
{
    Shape s;
    gf_status st = check_args(f, path, lat, Uv, s);
    if (st != GF_OK) return st;
    const HostArray U{const_cast<void *>(Uv), lat->dtype == GF_F64};
    const size_t n = (size_t)(72 * s.V);
    std::vector<unsigned char> out;
    switch (f->type) {
    case GF_FILE_PLAIN:
    case GF_FILE_HEADERONLY:
    case GF_FILE_VOGT: {
        const size_t rb = real_bytes(f, lat);
        if (f->type == GF_FILE_VOGT) out = f->have_header ? f->header : default_vogt_header(s, rb);
        if (f->type == GF_FILE_HEADERONLY && f->have_header) out = f->header;
        const size_t off = out.size();
        out.resize(off + n * rb);
        encode_standard(f, s, U, rb, out.data() + off);
        break;
    }
    case GF_FILE_ILDG: {
        std::vector<unsigned char> bin(8 * n);
        encode_ildg(f, s, U, bin.data());
        if (!f->have_header) {
            LimeRecord fmt, data;
            const std::string xml = ildg_format_xml(s);
            fmt.type = "ildg-format";
            fmt.mb = true;
            fmt.data.assign(xml.begin(), xml.end());
            data.type = kIldgBinary;
            data.me = true;
            data.data = std::move(bin);
            append_lime(out, fmt);
            append_lime(out, data);
            break;
        }
        for (const auto &r : f->records) {
            LimeRecord w;
            w.type = r.type;
            w.mb = r.mb;
            w.me = r.me;
            if (r.type == kIldgBinary) {
                w.data = bin;
            } else {
                // the text up to its first NUL; xlf-info gets the annealing steps
                std::string text((const char *)r.data.data(), strnlen((const char *)r.data.data(), r.data.size()));
                if (r.type == "xlf-info") text += " SA steps " + std::to_string(sa_steps);
                w.data.assign(text.begin(), text.end());
            }
            append_lime(out, w);
        }
        break;
    }
    case GF_FILE_QCDSTAG: {
        out.resize(8 * n);
        for_each_link_entry(s, [&](int64_t t, int64_t x, int64_t y, int64_t z, int mu, int i, int j, int c) {
            const double v = U.get(mem_index(s, f->pattern, t, x, y, z, mu, i, j, c));
            std::memcpy(out.data() + 8 * qcdstag_element(s, t, x, y, z, mu, i, j, c), &v, 8);
        });
        break;
    }
    default:
        return io_fail(GF_ERR_BAD_ARG, "unknown file type");
    }
    return write_all(path, out);
}

}  // extern "C"
