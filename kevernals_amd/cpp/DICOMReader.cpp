// DICOM ingest of projection images (DICOMReader.cpp of the reference, over
// a DICOM parser of its own: PS3.5 data encoding, PS3.10 file format).
#include "DICOMReader.h"

#include <omp.h>
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

bool read_all(const std::string &path, std::vector<uint8_t> *out)
{
    std::ifstream f(path, std::ios::binary);
    if (!f)
        return false;
    f.seekg(0, std::ios::end);
    const std::streamoff n = f.tellg();
    if (n < 0)
        return false;
    f.seekg(0);
    out->resize((size_t)n);
    return n == 0 || (bool)f.read(reinterpret_cast<char *>(out->data()), n);
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
bool locate(const std::vector<uint8_t> &b, size_t *start)
{
    if (b.size() >= 132 && std::memcmp(b.data() + 128, "DICM", 4) == 0) {
        *start = 132;
        return true;
    }
    if (b.size() >= 8) {
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
    return locate(head, &start);
}

bool ReadFile(const std::string &path, Frames *out, std::string *err)
{
    auto fail = [&](const std::string &m) {
        if (err)
            *err = path + ": " + m;
        return false;
    };
    std::vector<uint8_t> bytes;
    if (!read_all(path, &bytes))
        return fail("cannot read the file");
    size_t start;
    if (!locate(bytes, &start))
        return fail("not a DICOM file");

    // File meta information (group 0002): explicit VR little endian.
    Syntax syntax;
    syntax.explicitVR = true;
    {
        Parser meta(bytes.data(), bytes.size(), Syntax{});
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
            const bool vr = bytes.size() >= start + 6 && bytes[start + 4] >= 'A' && bytes[start + 4] <= 'Z' &&
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
    std::vector<uint8_t> inflated;
    const uint8_t *ds = bytes.data() + start;
    size_t dsn = bytes.size() - start;
    if (syntax.deflated) {
        if (!inflate_raw(ds, dsn, &inflated))
            return fail("deflated dataset does not inflate");
        ds = inflated.data();
        dsn = inflated.size();
    }

    Parser p(ds, dsn, syntax);
    int rows = -1, cols = -1, frames = 1, spp = 1, alloc = -1, stored = -1, high = -1, rep = 0;
    double pixelSpacing[2] = {0, 0}, imagerSpacing[2] = {0, 0};
    bool hasPixelSpacing = false, hasImagerSpacing = false;
    Element pixels;
    bool havePixels = false;
    Frames f;
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
        case 0x00100020: f.hasPatientID = true; break;
        case 0x00080032: if (e.length != kUndefined) f.acquisitionTime = text(p, e); break;
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

    f.rows = rows;
    f.columns = cols;
    f.frames = frames;
    const double *sp = hasPixelSpacing ? pixelSpacing : hasImagerSpacing ? imagerSpacing : nullptr;
    if (sp) {
        f.spacing[0] = sp[1];  // column spacing: x
        f.spacing[1] = sp[0];  // row spacing: y
    }
    f.values.resize(count);
    const uint8_t *px = ds + pixels.value;
    const int shift = high + 1 - stored;
    const uint64_t mask = stored == 32 ? 0xFFFFFFFFull : ((1ull << stored) - 1);
    const bool big = syntax.bigEndian;
#pragma omp parallel for if (count >= (1u << 20) && !omp_in_parallel()) schedule(static)
    for (size_t k = 0; k < count; ++k) {
        uint32_t raw;
        const uint8_t *q = px + k * bpp;
        if (bpp == 1)
            raw = q[0];
        else if (bpp == 2)
            raw = big ? (uint32_t)(q[0] << 8 | q[1]) : (uint32_t)(q[0] | q[1] << 8);
        else
            raw = big ? (uint32_t)q[0] << 24 | (uint32_t)q[1] << 16 | (uint32_t)q[2] << 8 | q[3]
                      : (uint32_t)q[0] | (uint32_t)q[1] << 8 | (uint32_t)q[2] << 16 | (uint32_t)q[3] << 24;
        const uint64_t v = ((uint64_t)raw >> shift) & mask;
        if (rep == 1 && (v >> (stored - 1)) & 1)
            f.values[k] = (float)((int64_t)v - ((int64_t)1 << stored));
        else
            f.values[k] = (float)v;
    }
    *out = std::move(f);
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
    if (!kvdicom::CanReadFile(p_dicomFilePath))
        return reading_problem("Cannot read file: " + p_dicomFilePath);
    if (!m_tomoGeometry)
        return reading_problem("DICOMReader: no geometry");
    kvdicom::Frames f;
    std::string err;
    if (!kvdicom::ReadFile(p_dicomFilePath, &f, &err))
        return reading_problem("Error reading file " + err);
    const Pixel bl = m_tomoGeometry->GetProjectionROIsBLPixelPositionOnDetector();
    const Size2D size = m_tomoGeometry->projectionsRoisSize();
    auto img = kvdicom::CropToImage(f, bl.x, bl.y, size.x, size.y, &err);
    if (img == nullptr)
        return reading_problem("Empty data for raw projections " + p_dicomFilePath + ": " + err);
    return img;
}

namespace kvdicom {
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
    struct Entry {
        Frames frames;
        std::string time, path, error;
        bool ok = false;
    };
    // files decode independently
    std::vector<Entry> images(paths.size());
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < (int)paths.size(); ++i) {
        Entry &e = images[i];
        e.path = paths[i];
        if (!CanReadFile(e.path)) {
            e.error = "Cannot read file: " + e.path;
            continue;
        }
        e.ok = ReadFile(e.path, &e.frames, &e.error);
        if (e.ok && e.frames.hasPatientID)  // DICOMReader.cpp:148-152
            e.time = e.frames.acquisitionTime;
    }
    for (const Entry &e : images)
        if (!e.ok)
            return fail(e.error);
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
    // vtkExtractVOI of each image (clipped to it), appended along z
    // (vtkImageAppend): every crop must cover the same x and y extent
    std::vector<int> zoff(images.size() + 1, 0);
    int ext[4] = {0, 0, 0, 0};
    for (size_t i = 0; i < images.size(); ++i) {
        const Frames &f = images[i].frames;
        const int e[4] = {std::max(x0, 0), std::min(x0 + w - 1, f.columns - 1), std::max(y0, 0),
                          std::min(y0 + h - 1, f.rows - 1)};
        if (w <= 0 || h <= 0 || e[0] > e[1] || e[2] > e[3])
            return fail("Empty data for raw projections " + images[i].path +
                        ": the detector region lies outside the image");
        if (i == 0)
            std::copy(e, e + 4, ext);
        else if (!std::equal(e, e + 4, ext))
            return fail("image " + images[i].path + " differs in size from the first image");
        if (verbose)
            std::cout << "Ordered raw projections " << images[i].path << std::endl;
        zoff[i + 1] = zoff[i] + f.frames;
    }
    auto img = ImageDataPtr::New();
    img->SetExtent(ext[0], ext[1], ext[2], ext[3], 0, zoff.back() - 1);
    img->SetSpacing(images[0].frames.spacing[0], images[0].frames.spacing[1], 1.0);
    img->AllocateScalars(VTK_FLOAT, 1);
    float *dst = static_cast<float *>(img->GetScalarPointer());
    const size_t nx = (size_t)(ext[1] - ext[0] + 1), ny = (size_t)(ext[3] - ext[2] + 1);
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < (int)images.size(); ++i) {
        const Frames &f = images[i].frames;
        for (int z = 0; z < f.frames; ++z)
            for (int y = ext[2]; y <= ext[3]; ++y) {
                const float *src = f.values.data() + ((size_t)z * f.rows + (f.rows - 1 - y)) * f.columns + ext[0];
                std::copy(src, src + nx, dst + (((size_t)zoff[i] + z) * ny + (size_t)(y - ext[2])) * nx);
            }
    }
    *out = img;
    return true;
}
}  // namespace kvdicom

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
        kvdicom::Frames f;
        if (!kvdicom::CanReadFile(path))
            return report(std::string("cannot read file: ") + path);
        if (!kvdicom::ReadFile(path, &f, &msg))
            return report(msg);
        auto img = roi ? kvdicom::CropToImage(f, roi[0], roi[1], roi[2], roi[3], &msg)
                       : kvdicom::CropToImage(f, 0, 0, f.columns, f.rows, &msg);
        if (img == nullptr)
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
