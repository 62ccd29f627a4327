// DICOM ingest of projection images (DICOMReader.cpp of the reference, over
// a DICOM parser of its own: PS3.5 data encoding, PS3.10 file format).
#include "DICOMReader.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include <zlib.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <iterator>
#include <memory>

#include "kevernals_dropin.h"

namespace {
struct DicomCategory : std::error_category {
    const char *name() const noexcept override { return "DICOMReader"; }
    std::string message(int) const override { return "DataFileReadingProblem"; }
};
const DicomCategory g_dicom_category;
}  // namespace

std::error_code make_error_code(DICOMReaderErrorCode e) { return {(int)e, g_dicom_category}; }

namespace kvdicom {
namespace {

constexpr uint32_t kUndefined = 0xFFFFFFFFu;

struct Syntax {
    bool explicitVR = true, bigEndian = false, deflated = false;
};

// One element header: tag, VR (explicit syntaxes), value offset and length.
struct Element {
    uint16_t group = 0, element = 0;
    char vr[2] = {0, 0};
    uint32_t length = 0;
    size_t value = 0;
};

class Parser {
public:
    Parser(const uint8_t *p, size_t n, Syntax s) : p_(p), n_(n), s_(s) {}

    size_t pos() const { return i_; }
    void seek(size_t i) { i_ = i; }
    bool done() const { return i_ >= n_; }
    const std::string &error() const { return err_; }
    const uint8_t *data() const { return p_; }
    size_t size() const { return n_; }
    const Syntax &syntax() const { return s_; }

    uint16_t u16(size_t at) const
    {
        return s_.bigEndian ? (uint16_t)(p_[at] << 8 | p_[at + 1]) : (uint16_t)(p_[at] | p_[at + 1] << 8);
    }
    uint32_t u32(size_t at) const
    {
        return s_.bigEndian ? (uint32_t)p_[at] << 24 | (uint32_t)p_[at + 1] << 16 | (uint32_t)p_[at + 2] << 8 | p_[at + 3]
                            : (uint32_t)p_[at] | (uint32_t)p_[at + 1] << 8 | (uint32_t)p_[at + 2] << 16 |
                                  (uint32_t)p_[at + 3] << 24;
    }

    // Reads the element header at the cursor and leaves the cursor on its value.
    bool header(Element *e)
    {
        if (i_ + 8 > n_)
            return fail("truncated element header");
        e->group = u16(i_);
        e->element = u16(i_ + 2);
        if (e->group == 0xFFFE) {  // items and delimiters carry no VR
            e->vr[0] = e->vr[1] = 0;
            e->length = u32(i_ + 4);
            i_ += 8;
        } else if (s_.explicitVR) {
            e->vr[0] = (char)p_[i_ + 4];
            e->vr[1] = (char)p_[i_ + 5];
            if (long_vr(e->vr)) {
                if (i_ + 12 > n_)
                    return fail("truncated element header");
                e->length = u32(i_ + 8);
                i_ += 12;
            } else {
                e->length = u16(i_ + 6);
                i_ += 8;
            }
        } else {
            e->vr[0] = e->vr[1] = 0;
            e->length = u32(i_ + 4);
            i_ += 8;
        }
        e->value = i_;
        if (e->length != kUndefined && e->length > n_ - i_)
            return fail("element value runs past the end of the file");
        return true;
    }

    // Skips the value of e (undefined lengths: a sequence of items).
    bool skip(const Element &e)
    {
        if (e.length != kUndefined) {
            i_ = e.value + e.length;
            return true;
        }
        if (e.vr[0] != 'U' || e.vr[1] != 'N')
            return skip_items(0);
        // UN of undefined length: items in implicit VR little endian (PS3.5 6.2.2)
        const Syntax saved = s_;
        s_.explicitVR = false;
        s_.bigEndian = false;
        const bool ok = skip_items(0);
        s_ = saved;
        return ok;
    }

    bool fail(const std::string &m)
    {
        if (err_.empty())
            err_ = m;
        return false;
    }

private:
    static bool long_vr(const char vr[2])
    {
        static const char *kLong[] = {"OB", "OD", "OF", "OL", "OV", "OW", "SQ", "SV", "UC", "UN", "UR", "UT", "UV"};
        for (const char *l : kLong)
            if (vr[0] == l[0] && vr[1] == l[1])
                return true;
        return false;
    }

    // Items up to the sequence delimiter; an item of undefined length holds
    // elements up to the item delimiter.
    bool skip_items(int depth)
    {
        if (depth > 64)
            return fail("sequences nested too deeply");
        for (;;) {
            Element it;
            if (!header(&it))
                return false;
            if (it.group == 0xFFFE && it.element == 0xE0DD)
                return true;  // sequence delimitation
            if (it.group != 0xFFFE || it.element != 0xE000)
                return fail("malformed sequence");
            if (it.length != kUndefined) {
                i_ = it.value + it.length;
                continue;
            }
            for (;;) {
                Element e;
                if (!header(&e))
                    return false;
                if (e.group == 0xFFFE && e.element == 0xE00D)
                    break;  // item delimitation
                if (e.length != kUndefined)
                    i_ = e.value + e.length;
                else if (!skip_items(depth + 1))
                    return false;
            }
        }
    }

    const uint8_t *p_;
    size_t n_;
    Syntax s_;
    size_t i_ = 0;
    std::string err_;
};

std::string text(const Parser &ps, const Element &e)
{
    std::string s(reinterpret_cast<const char *>(ps.data() + e.value), e.length);
    const size_t bs = s.find('\\');  // first value of a multi-valued string
    if (bs != std::string::npos)
        s.resize(bs);
    while (!s.empty() && (s.back() == ' ' || s.back() == '\0'))
        s.pop_back();
    return s;
}

bool inflate_raw(const uint8_t *p, size_t n, std::vector<uint8_t> *out)
{
    z_stream z{};
    if (inflateInit2(&z, -15) != Z_OK)
        return false;
    z.next_in = const_cast<Bytef *>(p);
    z.avail_in = (uInt)n;
    out->clear();
    uint8_t buf[1 << 16];
    int rc;
    do {
        z.next_out = buf;
        z.avail_out = sizeof buf;
        rc = inflate(&z, Z_NO_FLUSH);
        if (rc != Z_OK && rc != Z_STREAM_END) {
            inflateEnd(&z);
            return false;
        }
        out->insert(out->end(), buf, buf + (sizeof buf - z.avail_out));
    } while (rc != Z_STREAM_END && (z.avail_in > 0 || z.avail_out == 0));
    inflateEnd(&z);
    return rc == Z_STREAM_END;
}

// Start of the dataset: after the Part 10 preamble, or 0 for a bare dataset.
bool locate(const uint8_t *b, size_t n, size_t *start)
{
    if (n >= 132 && std::memcmp(b + 128, "DICM", 4) == 0) {
        *start = 132;
        return true;
    }
    if (n >= 8) {
        const uint16_t g = (uint16_t)(b[0] | b[1] << 8);
        if (g == 0x0002 || g == 0x0008) {
            *start = 0;
            return true;
        }
    }
    return false;
}

}  // namespace

bool CanReadFile(const std::string &path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f)
        return false;
    std::vector<uint8_t> head(132, 0);
    f.read(reinterpret_cast<char *>(head.data()), 132);
    head.resize((size_t)f.gcount());
    size_t start;
    return locate(head.data(), head.size(), &start);
}

namespace {

// A file's bytes, memory-mapped read-only (no copy, no fresh memory).
struct Mapped {
    const uint8_t *p = nullptr;
    size_t n = 0;
    void *map = nullptr;
    Mapped() = default;
    Mapped(const Mapped &) = delete;
    Mapped &operator=(const Mapped &) = delete;
    ~Mapped()
    {
        if (map)
            munmap(map, n);
    }
    bool open(const std::string &path)
    {
        const int fd = ::open(path.c_str(), O_RDONLY);
        if (fd < 0)
            return false;
        struct stat st;
        const bool ok = fstat(fd, &st) == 0 && st.st_size >= 0;
        if (ok && st.st_size > 0) {
            map = mmap(nullptr, (size_t)st.st_size, PROT_READ, MAP_PRIVATE, fd, 0);
            if (map == MAP_FAILED) {
                map = nullptr;
            } else {
                p = static_cast<const uint8_t *>(map);
                n = (size_t)st.st_size;
            }
        }
        ::close(fd);
        return ok && (st.st_size == 0 || map);
    }
};

// One file parsed up to its pixel cells (not decoded).
struct Parsed {
    Mapped file;
    std::vector<uint8_t> inflated;  // the dataset of a deflated file
    const uint8_t *cells = nullptr;
    bool big = false;
    int rows = 0, cols = 0, frames = 0, bpp = 0, stored = 0, shift = 0, rep = 0;
    double spacing[2] = {1.0, 1.0};
    bool hasPatientID = false;
    std::string acquisitionTime;
};

bool Parse(const std::string &path, Parsed *f, std::string *err)
{
    auto fail = [&](const std::string &m) {
        if (err)
            *err = path + ": " + m;
        return false;
    };
    if (!f->file.open(path))
        return fail("cannot read the file");
    const uint8_t *bytes = f->file.p;
    const size_t nbytes = f->file.n;
    size_t start;
    if (!locate(bytes, nbytes, &start))
        return fail("not a DICOM file");

    // File meta information (group 0002): explicit VR little endian.
    Syntax syntax;
    syntax.explicitVR = true;
    {
        Parser meta(bytes, nbytes, Syntax{});
        meta.seek(start);
        std::string ts;
        while (meta.pos() + 2 <= meta.size() && meta.u16(meta.pos()) == 0x0002) {
            Element e;
            if (!meta.header(&e))
                return fail(meta.error());
            if (e.element == 0x0010)
                ts = text(meta, e);
            if (!meta.skip(e))
                return fail(meta.error());
        }
        start = meta.pos();
        if (ts.empty()) {
            // bare dataset: implicit VR little endian unless a VR is present
            const bool vr = nbytes >= start + 6 && bytes[start + 4] >= 'A' && bytes[start + 4] <= 'Z' &&
                            bytes[start + 5] >= 'A' && bytes[start + 5] <= 'Z';
            syntax.explicitVR = vr;
        } else if (ts == "1.2.840.10008.1.2") {
            syntax.explicitVR = false;
        } else if (ts == "1.2.840.10008.1.2.1") {
        } else if (ts == "1.2.840.10008.1.2.1.99") {
            syntax.deflated = true;
        } else if (ts == "1.2.840.10008.1.2.2") {
            syntax.bigEndian = true;
        } else {
            return fail("transfer syntax " + ts + " (encapsulated pixel data) is not supported");
        }
    }
    const uint8_t *ds = bytes + start;
    size_t dsn = nbytes - start;
    if (syntax.deflated) {
        if (!inflate_raw(ds, dsn, &f->inflated))
            return fail("deflated dataset does not inflate");
        ds = f->inflated.data();
        dsn = f->inflated.size();
    }

    Parser p(ds, dsn, syntax);
    int rows = -1, cols = -1, frames = 1, spp = 1, alloc = -1, stored = -1, high = -1, rep = 0;
    double pixelSpacing[2] = {0, 0}, imagerSpacing[2] = {0, 0};
    bool hasPixelSpacing = false, hasImagerSpacing = false;
    Element pixels;
    bool havePixels = false;
    auto us = [&](const Element &e) { return e.length >= 2 ? (int)p.u16(e.value) : -1; };
    auto ds2 = [&](const Element &e, double v[2]) {
        const std::string s(reinterpret_cast<const char *>(ds + e.value), e.length);
        return std::sscanf(s.c_str(), " %lf \\ %lf", &v[0], &v[1]) == 2;
    };
    while (!p.done() && !havePixels) {
        Element e;
        if (!p.header(&e))
            return fail(p.error());
        const uint32_t tag = (uint32_t)e.group << 16 | e.element;
        switch (tag) {
        case 0x00100020: f->hasPatientID = true; break;
        case 0x00080032: if (e.length != kUndefined) f->acquisitionTime = text(p, e); break;
        case 0x00280002: spp = us(e); break;
        case 0x00280008: if (e.length != kUndefined) frames = std::atoi(text(p, e).c_str()); break;
        case 0x00280010: rows = us(e); break;
        case 0x00280011: cols = us(e); break;
        case 0x00280030: hasPixelSpacing = e.length != kUndefined && ds2(e, pixelSpacing); break;
        case 0x00181164: hasImagerSpacing = e.length != kUndefined && ds2(e, imagerSpacing); break;
        case 0x00280100: alloc = us(e); break;
        case 0x00280101: stored = us(e); break;
        case 0x00280102: high = us(e); break;
        case 0x00280103: rep = us(e); break;
        case 0x7FE00010:
            if (e.length == kUndefined)
                return fail("encapsulated (compressed) pixel data is not supported");
            pixels = e;
            havePixels = true;
            break;
        default: break;
        }
        if (!havePixels && !p.skip(e))
            return fail(p.error());
    }
    if (!havePixels)
        return fail("no pixel data");
    if (rows <= 0 || cols <= 0)
        return fail("missing or empty Rows / Columns");
    if (spp != 1)
        return fail("one sample per pixel is supported, the file has " + std::to_string(spp));
    if (alloc != 8 && alloc != 16 && alloc != 32)
        return fail("BitsAllocated must be 8, 16 or 32");
    if (frames <= 0)
        return fail("NumberOfFrames must be positive");
    if (stored < 0)
        stored = alloc;
    if (high < 0)
        high = stored - 1;
    if (stored < 1 || stored > alloc || high >= alloc || high + 1 < stored)
        return fail("inconsistent BitsStored / HighBit");
    const size_t bpp = (size_t)alloc / 8, count = (size_t)rows * cols * frames;
    if ((size_t)pixels.length / bpp < count)  // (no overflow for any header values)
        return fail("pixel data shorter than Rows x Columns x NumberOfFrames");
    f->cells = ds + pixels.value;
    f->big = syntax.bigEndian;
    f->rows = rows;
    f->cols = cols;
    f->frames = frames;
    f->bpp = (int)bpp;
    f->stored = stored;
    f->shift = high + 1 - stored;
    f->rep = rep;
    const double *sp = hasPixelSpacing ? pixelSpacing : hasImagerSpacing ? imagerSpacing : nullptr;
    if (sp) {
        f->spacing[0] = sp[1];  // column spacing: x
        f->spacing[1] = sp[0];  // row spacing: y
    }
    return true;
}

// Stored values of nx cells of one file row (first file row = 0) from
// column x0: the BitsStored bits ending at HighBit, sign-extended when
// PixelRepresentation is 1.
void decode_row(const Parsed &f, int frame, int file_row, int x0, int nx, float *dst)
{
    const uint8_t *q = f.cells + (((size_t)frame * f.rows + file_row) * f.cols + x0) * f.bpp;
    const uint64_t mask = f.stored == 32 ? 0xFFFFFFFFull : ((1ull << f.stored) - 1);
    const int64_t sign = f.rep == 1 ? (int64_t)1 << (f.stored - 1) : 0;
    auto put = [&](int i, uint32_t raw) {
        const uint64_t v = ((uint64_t)raw >> f.shift) & mask;
        dst[i] = sign && (v & (uint64_t)sign) ? (float)((int64_t)v - 2 * sign) : (float)v;
    };
    if (f.bpp == 1) {
        for (int i = 0; i < nx; ++i)
            put(i, q[i]);
    } else if (f.bpp == 2) {
        for (int i = 0; i < nx; ++i, q += 2)
            put(i, f.big ? (uint32_t)(q[0] << 8 | q[1]) : (uint32_t)(q[0] | q[1] << 8));
    } else {
        for (int i = 0; i < nx; ++i, q += 4)
            put(i, f.big ? (uint32_t)q[0] << 24 | (uint32_t)q[1] << 16 | (uint32_t)q[2] << 8 | q[3]
                         : (uint32_t)q[0] | (uint32_t)q[1] << 8 | (uint32_t)q[2] << 16 | (uint32_t)q[3] << 24);
    }
}

// vtkExtractVOI's extent: the requested region clipped to the image.
bool crop_extent(const Parsed &f, int x0, int y0, int w, int h, int e[4])
{
    e[0] = std::max(x0, 0), e[1] = std::min(x0 + w - 1, f.cols - 1);
    e[2] = std::max(y0, 0), e[3] = std::min(y0 + h - 1, f.rows - 1);
    return w > 0 && h > 0 && e[0] <= e[1] && e[2] <= e[3];
}

// Decodes each file's frames, cropped and bottom-up, into z-slabs of out:
// file i fills slices zoff[i] .. zoff[i+1]-1 of an [z][ny][nx] image.
void decode_into(const std::vector<const Parsed *> &files, const std::vector<int> &zoff, const int e[4], float *out)
{
    const int nx = e[1] - e[0] + 1, ny = e[3] - e[2] + 1;
    const long long rows = (long long)zoff.back() * ny;
    std::vector<int> file_of(zoff.back());
    for (size_t i = 0; i < files.size(); ++i)
        for (int z = zoff[i]; z < zoff[i + 1]; ++z)
            file_of[z] = (int)i;
#pragma omp parallel for schedule(static)
    for (long long r = 0; r < rows; ++r) {
        const int z = (int)(r / ny), y = e[2] + (int)(r % ny);
        const Parsed &f = *files[file_of[z]];
        decode_row(f, z - zoff[file_of[z]], f.rows - 1 - y, e[0], nx, out + (size_t)r * nx);
    }
}

}  // namespace

bool ReadFile(const std::string &path, Frames *out, std::string *err)
{
    Parsed f;
    if (!Parse(path, &f, err))
        return false;
    Frames r;
    r.rows = f.rows;
    r.columns = f.cols;
    r.frames = f.frames;
    r.spacing[0] = f.spacing[0];
    r.spacing[1] = f.spacing[1];
    r.hasPatientID = f.hasPatientID;
    r.acquisitionTime = f.acquisitionTime;
    r.values.resize((size_t)f.rows * f.cols * f.frames);
    for (int z = 0; z < f.frames; ++z)
        for (int y = 0; y < f.rows; ++y)
            decode_row(f, z, y, 0, f.cols, r.values.data() + ((size_t)z * f.rows + y) * f.cols);
    *out = std::move(r);
    return true;
}

ImageDataPtr CropToImage(const Frames &f, int x0, int y0, int w, int h, std::string *err)
{
    // vtkExtractVOI: the requested extent clipped to the image's
    const int xa = std::max(x0, 0), xb = std::min(x0 + w - 1, f.columns - 1);
    const int ya = std::max(y0, 0), yb = std::min(y0 + h - 1, f.rows - 1);
    if (w <= 0 || h <= 0 || xa > xb || ya > yb) {
        if (err)
            *err = "the detector region lies outside the image";
        return nullptr;
    }
    auto img = ImageDataPtr::New();
    img->SetExtent(xa, xb, ya, yb, 0, f.frames - 1);
    img->SetSpacing(f.spacing[0], f.spacing[1], 1.0);
    img->AllocateScalars(VTK_FLOAT, 1);
    float *out = static_cast<float *>(img->GetScalarPointer());
    const size_t nx = (size_t)(xb - xa + 1);
    for (int z = 0; z < f.frames; ++z)
        for (int y = ya; y <= yb; ++y) {
            const int fileRow = f.rows - 1 - y;  // bottom-up rows
            const float *src = f.values.data() + ((size_t)z * f.rows + fileRow) * f.columns + xa;
            std::copy(src, src + nx, out + ((size_t)z * (yb - ya + 1) + (y - ya)) * nx);
        }
    return img;
}

bool ReadImage(const std::string &path, int x0, int y0, int w, int h, ImageDataPtr *out, std::string *err)
{
    Parsed f;
    if (!CanReadFile(path)) {
        if (err)
            *err = "Cannot read file: " + path;
        return false;
    }
    if (!Parse(path, &f, err))
        return false;
    int e[4];
    if (!crop_extent(f, x0, y0, w, h, e)) {
        if (err)
            *err = "Empty data for raw projections " + path + ": the detector region lies outside the image";
        return false;
    }
    auto img = ImageDataPtr::New();
    img->SetExtent(e[0], e[1], e[2], e[3], 0, f.frames - 1);
    img->SetSpacing(f.spacing[0], f.spacing[1], 1.0);
    img->AllocateScalars(VTK_FLOAT, 1);
    decode_into({&f}, {0, f.frames}, e, static_cast<float *>(img->GetScalarPointer()));
    *out = img;
    return true;
}

bool ReadDirectoryImages(const std::string &dir, const std::vector<int> &indicesToRemove, int x0, int y0, int w,
                         int h, bool verbose, ImageDataPtr *out, std::string *err)
{
    namespace fs = std::filesystem;
    auto fail = [&](const std::string &m) {
        if (err)
            *err = m;
        return false;
    };
    std::vector<std::string> paths;
    std::error_code ec;
    for (const auto &entry : fs::directory_iterator(dir, ec)) {
        const std::string ext = entry.path().extension().generic_string();
        if (ext == ".dcm" || ext == ".DCM")
            paths.push_back(entry.path().generic_string());
    }
    if (ec)
        return fail("cannot list " + dir + ": " + ec.message());
    std::sort(paths.begin(), paths.end());
    // headers parse independently (the pixel cells stay mapped, undecoded)
    std::vector<Parsed> parsed(paths.size());
    std::vector<std::string> errors(paths.size());
    std::vector<char> ok(paths.size(), 0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < (int)paths.size(); ++i) {
        if (!CanReadFile(paths[i]))
            errors[i] = "Cannot read file: " + paths[i];
        else
            ok[i] = Parse(paths[i], &parsed[i], &errors[i]);
    }
    for (size_t i = 0; i < paths.size(); ++i)
        if (!ok[i])
            return fail(errors[i]);
    struct Entry {
        const Parsed *f;
        std::string time, path;
    };
    std::vector<Entry> images;
    for (size_t i = 0; i < paths.size(); ++i)  // DICOMReader.cpp:148-152
        images.push_back({&parsed[i], parsed[i].hasPatientID ? parsed[i].acquisitionTime : std::string(), paths[i]});
    std::stable_sort(images.begin(), images.end(), [](const Entry &a, const Entry &b) { return a.time < b.time; });
    std::vector<int> drop = indicesToRemove;
    std::sort(drop.begin(), drop.end(), std::greater<int>());
    for (int index : drop) {
        if (index < 0 || index >= (int)images.size())
            return fail("index to remove " + std::to_string(index) + " is out of range");
        images.erase(images.begin() + index);
    }
    if (images.empty())
        return fail("no DICOM images in " + dir);
    // vtkExtractVOI of each image, appended along z (vtkImageAppend): every
    // crop must cover the same x and y extent
    std::vector<int> zoff(images.size() + 1, 0);
    std::vector<const Parsed *> files;
    int ext[4] = {0, 0, 0, 0};
    for (size_t i = 0; i < images.size(); ++i) {
        int e[4];
        if (!crop_extent(*images[i].f, x0, y0, w, h, e))
            return fail("Empty data for raw projections " + images[i].path +
                        ": the detector region lies outside the image");
        if (i == 0)
            std::copy(e, e + 4, ext);
        else if (!std::equal(e, e + 4, ext))
            return fail("image " + images[i].path + " differs in size from the first image");
        if (verbose)
            std::cout << "Ordered raw projections " << images[i].path << std::endl;
        zoff[i + 1] = zoff[i] + images[i].f->frames;
        files.push_back(images[i].f);
    }
    auto img = ImageDataPtr::New();
    img->SetExtent(ext[0], ext[1], ext[2], ext[3], 0, zoff.back() - 1);
    img->SetSpacing(images[0].f->spacing[0], images[0].f->spacing[1], 1.0);
    img->AllocateScalars(VTK_FLOAT, 1);
    decode_into(files, zoff, ext, static_cast<float *>(img->GetScalarPointer()));
    *out = img;
    return true;
}
}  // namespace kvdicom

namespace {
Result<ImageDataPtr> reading_problem(const std::string &what)
{
    std::cout << what << std::endl;
    std::cout << "DataFileReadingProblem" << std::endl;
    return DICOMReaderErrorCode::DataFileReadingProblem;
}
}  // namespace

Result<ImageDataPtr> DICOMReader::Read(std::string const &p_dicomFilePath) const
{
    if (!std::filesystem::exists(p_dicomFilePath))
        return reading_problem("Raw projections dicom file does not exist: " + p_dicomFilePath);
    if (!m_tomoGeometry)
        return reading_problem("DICOMReader: no geometry");
    const Pixel bl = m_tomoGeometry->GetProjectionROIsBLPixelPositionOnDetector();
    const Size2D size = m_tomoGeometry->projectionsRoisSize();
    ImageDataPtr img;
    std::string err;
    if (!kvdicom::ReadImage(p_dicomFilePath, bl.x, bl.y, size.x, size.y, &img, &err))
        return reading_problem(err);
    return img;
}

Result<ImageDataPtr> DICOMReader::ReadDirectory(std::string const &p_dicomFilesContainedDirPath,
                                                std::vector<int> const &p_indicesToRemove) const
{
    if (!std::filesystem::exists(p_dicomFilesContainedDirPath))
        return reading_problem("Raw projections dicom files directory does not exist: " +
                               p_dicomFilesContainedDirPath);
    if (!m_tomoGeometry)
        return reading_problem("DICOMReader: no geometry");
    const Pixel bl = m_tomoGeometry->GetProjectionROIsBLPixelPositionOnDetector();
    const Size2D size = m_tomoGeometry->projectionsRoisSize();
    ImageDataPtr out;
    std::string err;
    if (!kvdicom::ReadDirectoryImages(p_dicomFilesContainedDirPath, p_indicesToRemove, bl.x, bl.y, size.x, size.y,
                                      true, &out, &err))
        return reading_problem(err);
    return out;
}

// ---- C entry points ------------------------------------------------------
struct kvd_image {
    ImageDataPtr img;
};

extern "C" kvd_image *kvd_dicom_read(const char *path, int directory, const int32_t *remove, int nremove,
                                     const int32_t *roi, char *err, int err_len)
{
    std::string msg;
    auto report = [&](const std::string &m) -> kvd_image * {
        if (err && err_len > 0)
            std::snprintf(err, (size_t)err_len, "%s", m.c_str());
        return nullptr;
    };
    if (!path)
        return report("no path");
    if (err && err_len > 0)
        err[0] = '\0';
    // the detector region: {x, y, width, height}, or the whole image
    if (!directory) {
        ImageDataPtr img;
        const bool ok = roi ? kvdicom::ReadImage(path, roi[0], roi[1], roi[2], roi[3], &img, &msg)
                            : kvdicom::ReadImage(path, 0, 0, 1 << 30, 1 << 30, &img, &msg);
        if (!ok)
            return report(msg);
        return new kvd_image{img};
    }
    int x0 = 0, y0 = 0, w = 1 << 30, h = 1 << 30;
    if (roi)
        x0 = roi[0], y0 = roi[1], w = roi[2], h = roi[3];
    std::vector<int> drop(remove, remove + std::max(nremove, 0));
    ImageDataPtr img;
    if (!kvdicom::ReadDirectoryImages(path, drop, x0, y0, w, h, false, &img, &msg))
        return report(msg);
    return new kvd_image{img};
}

extern "C" int kvd_image_info(const kvd_image *img, int32_t *extent, double *spacing)
{
    if (!img || img->img == nullptr)
        return 1;
    const int *e = img->img->GetExtent();
    const double *s = img->img->GetSpacing();
    for (int i = 0; i < 6 && extent; ++i)
        extent[i] = e[i];
    for (int i = 0; i < 3 && spacing; ++i)
        spacing[i] = s[i];
    return 0;
}

extern "C" const float *kvd_image_data(const kvd_image *img)
{
    return img && img->img != nullptr ? static_cast<const float *>(img->img->GetScalarPointer()) : nullptr;
}

extern "C" void kvd_image_free(kvd_image *img) { delete img; }
