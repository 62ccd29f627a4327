// TIFF writer with vtkTIFFWriter's layout (VTK 9.0.1 over libtiff 4.0.x).
#include "TiffWriter.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "kevernals_dropin.h"

namespace kvtiff {

size_t PackBitsRow(const uint8_t *bp, size_t cc, uint8_t *out)
{
    // The encoder's four states (tif_packbits.c): nothing pending, a literal
    // open at `lit`, a run just written, a run written after a literal.
    enum { Base, Literal, Run, LiteralRun } state = Base;
    uint8_t *op = out, *lit = nullptr;
    while (cc > 0) {
        const uint8_t b = *bp++;
        --cc;
        long n = 1;
        while (cc > 0 && *bp == b) {
            ++n;
            ++bp;
            --cc;
        }
        for (;;) {
            if (state == LiteralRun) {
                // a 2-byte run followed by a single byte joins the literal
                if (n == 1 && op[-2] == 0xFF && *lit < 126) {
                    *lit += 2;
                    state = *lit == 127 ? Base : Literal;
                    op[-2] = op[-1];
                } else {
                    state = Run;
                }
                continue;
            }
            if (n > 1) {
                state = state == Literal ? LiteralRun : Run;
                if (n > 128) {
                    *op++ = 0x81;  // -127: 128 copies
                    *op++ = b;
                    n -= 128;
                    continue;
                }
                *op++ = (uint8_t)(1 - n);
                *op++ = b;
            } else if (state == Literal) {
                if (++*lit == 127)
                    state = Base;
                *op++ = b;
            } else {
                lit = op;
                *op++ = 0;
                *op++ = b;
                state = Literal;
            }
            break;
        }
    }
    return (size_t)(op - out);
}

namespace {

enum : uint16_t { kShort = 3, kLong = 4, kRational = 5 };

struct Entry {
    uint16_t tag, type;
    std::vector<uint32_t> v;  // SHORT values, LONG values, or RATIONAL num/den pairs
    uint32_t count() const { return type == kRational ? (uint32_t)v.size() / 2 : (uint32_t)v.size(); }
    size_t bytes() const { return (type == kShort ? 2 : 4) * v.size(); }
    void pack(uint8_t *p) const
    {
        for (size_t i = 0; i < v.size(); ++i) {
            if (type == kShort) {
                const uint16_t s = (uint16_t)v[i];
                std::memcpy(p + 2 * i, &s, 2);
            } else {
                std::memcpy(p + 4 * i, &v[i], 4);
            }
        }
    }
};

// libtiff's data emission order within a directory (TIFFWriteDirectorySec).
const uint16_t kOrder[] = {256, 257, 282, 283, 254, 258, 259, 262, 274, 277, 278, 284, 296, 297, 279, 273, 338, 339};

// libtiff 4.0.x rational of a value stored as float.
void rational(double value, std::vector<uint32_t> &out)
{
    const double v = (double)(float)value;
    uint32_t num, den;
    if (!(v > 0.0)) {
        num = 0, den = 1;
    } else if (v < 4294967296.0 && v == (double)(uint32_t)v) {
        num = (uint32_t)v, den = 1;
    } else if (v < 1.0) {
        num = (uint32_t)(v * 4294967295.0), den = 0xFFFFFFFFu;
    } else {
        num = 0xFFFFFFFFu, den = std::isfinite(v) ? (uint32_t)(4294967295.0 / v) : 0u;
    }
    out = {num, den};
}

class File {
public:
    explicit File(const std::string &path) : f_(std::fopen(path.c_str(), "wb"))
    {
        const uint8_t header[8] = {'I', 'I', 42, 0, 0, 0, 0, 0};
        put(header, 8);
    }
    ~File()
    {
        if (f_)
            std::fclose(f_);
    }
    bool ok() const { return f_ && ok_; }
    uint64_t pos() const { return pos_; }
    void put(const void *p, size_t n)
    {
        if (!f_ || !ok_ || n == 0)
            return;
        ok_ = std::fwrite(p, 1, n, f_) == n;
        pos_ += n;
    }
    void patch32(uint64_t at, uint32_t value)
    {
        if (!ok())
            return;
        ok_ = std::fseek(f_, (long)at, SEEK_SET) == 0 && std::fwrite(&value, 4, 1, f_) == 1 &&
              std::fseek(f_, 0, SEEK_END) == 0;
    }
    bool close()
    {
        const bool good = ok() && std::fclose(f_) == 0;
        f_ = nullptr;
        return good;
    }

    // One directory at the even-rounded end of the file, linked from the
    // previous one (or the header).
    void directory(std::vector<Entry> entries)
    {
        if (pos_ & 1) {
            const uint8_t z = 0;
            put(&z, 1);
        }
        const uint64_t diroff = pos_;
        std::sort(entries.begin(), entries.end(), [](const Entry &a, const Entry &b) { return a.tag < b.tag; });
        const size_t n = entries.size();
        uint64_t dataoff = diroff + 2 + 12 * n + 4;
        std::vector<uint8_t> ifd(2 + 12 * n + 4, 0), data;
        const uint16_t nd = (uint16_t)n;
        std::memcpy(ifd.data(), &nd, 2);
        std::vector<uint32_t> where(n, 0);
        for (uint16_t tag : kOrder) {
            for (size_t i = 0; i < n; ++i) {
                if (entries[i].tag != tag || entries[i].bytes() <= 4)
                    continue;
                where[i] = (uint32_t)(dataoff + data.size());
                const size_t at = data.size();
                data.resize(at + entries[i].bytes());
                entries[i].pack(data.data() + at);
                if (data.size() & 1)
                    data.push_back(0);
            }
        }
        for (size_t i = 0; i < n; ++i) {
            uint8_t *e = ifd.data() + 2 + 12 * i;
            const uint32_t cnt = entries[i].count();
            std::memcpy(e, &entries[i].tag, 2);
            std::memcpy(e + 2, &entries[i].type, 2);
            std::memcpy(e + 4, &cnt, 4);
            if (entries[i].bytes() <= 4)
                entries[i].pack(e + 8);
            else
                std::memcpy(e + 8, &where[i], 4);
        }
        put(ifd.data(), ifd.size());
        put(data.data(), data.size());
        patch32(last_next_, (uint32_t)diroff);
        last_next_ = diroff + 2 + 12 * n;
        (void)dataoff;
    }

private:
    std::FILE *f_;
    bool ok_ = true;
    uint64_t pos_ = 0;
    uint64_t last_next_ = 4;
};

// One page: strips of whole rows, then its directory.
void page(File &f, const uint8_t *const *rows, int nx, int ny, int spp, int bytes_per_sample, bool is_float,
          bool has_res, double rx, double ry, int compression, std::vector<Entry> extra)
{
    const size_t line = (size_t)nx * spp * bytes_per_sample;
    const uint32_t rps = (uint32_t)std::max<size_t>(1, 8192 / std::max<size_t>(line, 1));
    const int nstrips = (int)((ny + rps - 1) / rps);
    std::vector<uint32_t> offsets(nstrips), counts(nstrips);
    // encode a batch of strips in parallel, write them in order
    const int batch = 256;
    std::vector<std::vector<uint8_t>> enc(batch);
    for (int s0 = 0; s0 < nstrips; s0 += batch) {
        const int s1 = std::min(nstrips, s0 + batch);
#pragma omp parallel for schedule(dynamic, 4)
        for (int s = s0; s < s1; ++s) {
            std::vector<uint8_t> &out = enc[s - s0];
            const int r0 = s * (int)rps, r1 = std::min(ny, r0 + (int)rps);
            out.resize(compression == kPackBits ? PackBitsBound(line) * (r1 - r0) : line * (r1 - r0));
            size_t used = 0;
            for (int r = r0; r < r1; ++r) {
                if (compression == kPackBits) {
                    used += PackBitsRow(rows[r], line, out.data() + used);
                } else {
                    std::memcpy(out.data() + used, rows[r], line);
                    used += line;
                }
            }
            out.resize(used);
        }
        for (int s = s0; s < s1; ++s) {
            offsets[s] = (uint32_t)f.pos();
            counts[s] = (uint32_t)enc[s - s0].size();
            f.put(enc[s - s0].data(), enc[s - s0].size());
        }
    }
    auto sl = [](uint16_t tag, uint32_t v) { return Entry{tag, v <= 0xFFFF ? kShort : kLong, {v}}; };
    std::vector<Entry> e = {
        sl(256, (uint32_t)nx),
        sl(257, (uint32_t)ny),
        Entry{258, kShort, std::vector<uint32_t>(spp, (uint32_t)(8 * bytes_per_sample))},
        Entry{259, kShort, {compression == kPackBits ? 32773u : 1u}},
        Entry{262, kShort, {spp == 1 ? 1u : 2u}},
        Entry{273, kLong, offsets},
        Entry{277, kShort, {(uint32_t)spp}},
        sl(278, rps),
        Entry{279, kLong, counts},
        Entry{284, kShort, {1}},
    };
    if (has_res) {
        Entry x{282, kRational, {}}, y{283, kRational, {}};
        rational(rx, x.v);
        rational(ry, y.v);
        e.push_back(x);
        e.push_back(y);
        e.push_back(Entry{296, kShort, {3}});  // centimetre
    }
    if (is_float)
        e.push_back(Entry{339, kShort, std::vector<uint32_t>(spp, 3u)});  // IEEE floating point
    for (auto &x : extra)
        e.push_back(std::move(x));
    f.directory(std::move(e));
}

}  // namespace

bool WriteImage(const std::string &path, const void *data, int scalar_type, int components, int nx, int ny, int nz,
                double spacing_x, double spacing_y, int compression, std::string *err)
{
    auto fail = [&](const std::string &m) {
        if (err)
            *err = m;
        return false;
    };
    int bps;
    switch (scalar_type) {
    case kUnsignedChar: bps = 1; break;
    case kUnsignedShort: bps = 2; break;
    case kFloat: bps = 4; break;
    default: return fail("TIFF output accepts unsigned char, unsigned short or float scalars");
    }
    if (!data || nx <= 0 || ny <= 0 || nz <= 0 || components <= 0)
        return fail("TIFF output: empty image");
    if (nx > 0xFFFF * 16 || ny > 0xFFFF * 16)
        return fail("TIFF output: image too large");
    if (nz > 1 && components != 1)
        return fail("TIFF output: a multi-page file holds one component per pixel");
    if (compression != kPackBits && compression != kNoCompression)
        return fail("TIFF output: compression must be none or PackBits");
    const size_t line = (size_t)nx * components * bps;
    if ((double)line * ny * nz > 3.9e9)
        return fail("TIFF output: more than 4 GiB does not fit a classic TIFF file");
    File f(path);
    if (!f.ok())
        return fail("cannot open " + path + " for writing");
    const double rx = 10.0 / spacing_x, ry = 10.0 / spacing_y;
    const bool has_res = rx > 0.0 && ry > 0.0;
    const uint8_t *base = static_cast<const uint8_t *>(data);
    std::vector<const uint8_t *> rows(ny);
    if (nz == 1) {
        for (int r = 0; r < ny; ++r)
            rows[r] = base + (size_t)(ny - 1 - r) * line;  // top row (highest y) first
        std::vector<Entry> extra = {Entry{274, kShort, {1}}};
        if (components > 3) {
            Entry es{338, kShort, std::vector<uint32_t>(components - 3, 0u)};
            es.v[0] = 1;  // associated alpha
            extra.push_back(es);
        }
        page(f, rows.data(), nx, ny, components, bps, scalar_type == kFloat, has_res, rx, ry, compression, extra);
    } else {
        for (int z = 0; z < nz; ++z) {
            for (int r = 0; r < ny; ++r)
                rows[r] = base + ((size_t)z * ny + r) * line;
            std::vector<Entry> extra = {Entry{254, kLong, {2}}, Entry{297, kShort, {(uint32_t)z, (uint32_t)nz}}};
            page(f, rows.data(), nx, ny, 1, bps, bps == 4, has_res, rx, ry, compression, extra);
            if (!f.ok())
                break;
        }
    }
    if (!f.close())
        return fail("I/O error writing " + path);
    return true;
}

}  // namespace kvtiff

#ifndef KV_USE_REAL_VTK
void vtkTIFFWriter::Write()
{
    error_ = 0;
    message_.clear();
    if (!input_ || !input_->GetScalarPointer()) {
        error_ = 1;
        message_ = "vtkTIFFWriter: no input data";
        return;
    }
    const int *d = input_->GetDimensions();
    const double *sp = input_->GetSpacing();
    std::string err;
    if (!kvtiff::WriteImage(file_, input_->GetScalarPointer(), kvtiff::kFloat, 1, d[0], d[1], d[2], sp[0], sp[1],
                            compression_ == PackBits ? kvtiff::kPackBits : kvtiff::kNoCompression, &err)) {
        error_ = 1;
        message_ = err;
    }
}
#endif

extern "C" int kvd_write_tiff(const char *path, const void *data, int scalar_type, int components, int nx, int ny,
                              int nz, double spacing_x, double spacing_y, int compression, char *err, int err_len)
{
    std::string msg;
    const bool ok = path && kvtiff::WriteImage(path, data, scalar_type, components, nx, ny, nz, spacing_x, spacing_y,
                                               compression, &msg);
    if (!path)
        msg = "no file name";
    if (err && err_len > 0) {
        std::snprintf(err, (size_t)err_len, "%s", ok ? "" : msg.c_str());
    }
    return ok ? 0 : 1;
}
