// Reading rules of the reference's XML acquisition-geometry file
// (KevLoq/Kevernals source/modules/geometry/TomoGeometry.cpp:13-437), on
// this project's XML reader (xml_lite.h).  The file lists, under its root
// element:
//   Rotation/center         the fulcrum, kept as written            (:39-59)
//   Grid/center             the frame origin subtracted from every source
//                           and detector position                  (:66-86)
//   Grid/scale, resolution  "a b c", taken in reverse order as the
//                           volume's (x, y, z) spacing and size; the
//                           size is read as floats and truncated    (:89-121)
//   XRay/sources/*          one source position per view           (:135-173)
//   Camera/totalWidth,totalHeight   the detector's size in mm, not ~0
//   Camera/pixelWidth,pixelHeight   its size in pixels (truncated), not 0
//                                                                   (:179-306)
//   Camera/references/*     per view the detector's bottom-left corner and
//                           height                                  (:307-339)
//   Radios/rois/*           "x0 y0 . . x1 y1" pixel corners of the region
//                           read, identical for every entry, clamped to the
//                           detector                                (:353-414)
// The pixel pitch is size / pixels in FP32 (PixelSpacing.cpp:6-25); the
// region's corner on view p is the detector corner plus the pitch summed
// x0 (y0) times in FP32 (TomoProjectionsSet.cpp:170-194, :137-140); the
// detector height is the mean of the per-view heights and the source x, z
// their means (TomoGeometry.cpp:342-344, TomoTable.cpp:4-19, applied by the
// value constructor of TomoGeometry).
//
// Numbers are parsed with std::stod as the reference does (a number prefix
// is read, "inf"/"nan" accepted) and narrowed the same way.  Where the
// reference's behaviour is undefined or it crashes (a missing Rotation/center
// text, a region entry with fewer than six words, integers out of range, an
// empty or inverted region, non-positive pixel counts, a comment inside a
// list), this reader reports an error instead.
#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <optional>
#include <sstream>
#include <stdexcept>

#include "Marshal.h"
#include "TomoGeometry.h"
#include "kevernals_dropin.h"
#include "xml_lite.h"

namespace {

// glob::SentenceToWords (commons/GlobalUtils.h:14-28)
std::vector<std::string> words_of(const std::string &s)
{
    std::istringstream in(s);
    std::vector<std::string> w;
    std::string t;
    while (in >> t)
        w.push_back(t);
    return w;
}

std::optional<double> parse_double(const std::string &s)
{
    try {
        return std::stod(s);
    } catch (const std::invalid_argument &) {
        return std::nullopt;
    } catch (const std::out_of_range &) {
        return std::nullopt;
    }
}

// Position3D::FromString (Dim3.h:71-95): exactly three words, each stod
// narrowed to float.
std::optional<Position3D> parse_position(const std::string &s)
{
    const auto w = words_of(s);
    if (w.size() != 3)
        return std::nullopt;
    const auto x = parse_double(w[0]), y = parse_double(w[1]), z = parse_double(w[2]);
    if (!x || !y || !z)
        return std::nullopt;
    return Position3D((float)*x, (float)*y, (float)*z);
}

bool to_int(double v, int &out)
{
    if (!(v > (double)INT_MIN - 1.0 && v < (double)INT_MAX + 1.0))
        return false;
    out = (int)v;  // truncation toward zero, as the implicit conversions
    return true;
}

struct Reader {
    TomoGeometryDescription d;

    bool fail(const std::string &msg)
    {
        d.ok = false;
        d.error = msg;
        return false;
    }
    const std::string *text_of(const xml_lite::Node *parent, const char *child, const char *path)
    {
        const xml_lite::Node *e = parent ? parent->first_child_element(child) : nullptr;
        if (!e) {
            fail(std::string("no \"") + path + "\" element");
            return nullptr;
        }
        const std::string *t = e->get_text();
        if (!t)
            fail(std::string("\"") + path + "\" element has no text");
        return t;
    }
    std::optional<Position3D> position(const xml_lite::Node *parent, const char *child, const char *path)
    {
        const std::string *t = text_of(parent, child, path);
        if (!t)
            return std::nullopt;
        auto p = parse_position(*t);
        if (!p)
            fail(std::string("\"") + path + "\" element could not be parsed");
        return p;
    }
    std::optional<float> number(const xml_lite::Node *parent, const char *child, const char *path)
    {
        const std::string *t = text_of(parent, child, path);
        if (!t)
            return std::nullopt;
        const auto v = parse_double(*t);
        if (!v) {
            fail(std::string("\"") + path + "\" element parsing problem: " + *t);
            return std::nullopt;
        }
        return (float)*v;
    }

    bool read(const std::string &path)
    {
        std::string err;
        auto root = xml_lite::load_file(path, err);
        if (!root)
            return fail("XML error: " + err);

        // fulcrum (:39-59)
        const xml_lite::Node *rotation = root->first_child_element("Rotation");
        if (!rotation)
            return fail("no \"Rotation\" element");
        const auto fulcrum = position(rotation, "center", "Rotation/center");
        if (!fulcrum)
            return false;
        d.fulcrum = *fulcrum;

        // reconstruction volume (:66-127)
        const xml_lite::Node *grid = root->first_child_element("Grid");
        if (!grid)
            return fail("no \"Grid\" element");
        const auto center = position(grid, "center", "Grid/center");
        if (!center)
            return false;
        const auto scale = position(grid, "scale", "Grid/scale");
        if (!scale)
            return false;
        const auto resolution = position(grid, "resolution", "Grid/resolution");
        if (!resolution)
            return false;
        d.voxelSpacing = VoxelSpacing(scale->z, scale->y, scale->x);
        int rx, ry, rz;
        if (!to_int(resolution->z, rx) || !to_int(resolution->y, ry) || !to_int(resolution->x, rz))
            return fail("\"Grid/resolution\" out of range");
        d.volumeSize = Size3D(rx, ry, rz);

        // sources, in the volume-centred frame (:135-173)
        const xml_lite::Node *xray = root->first_child_element("XRay");
        if (!xray)
            return fail("no \"XRay\" element");
        const xml_lite::Node *sources = xray->first_child_element("sources");
        if (!sources)
            return fail("no \"XRay/sources\" element");
        for (const xml_lite::Node *s : sources->child_elements()) {
            const std::string *t = s->get_text();
            const auto p = t ? parse_position(*t) : std::nullopt;
            if (!p)
                return fail("one among \"XRay/sources\" elements could not be parsed");
            d.sources.emplace_back(p->x - center->x, p->y - center->y, p->z - center->z);
        }
        if (d.sources.empty())
            return fail("\"XRay/sources\" is empty");

        // detector (:179-306)
        const xml_lite::Node *camera = root->first_child_element("Camera");
        if (!camera)
            return fail("no \"Camera\" element");
        const auto width = number(camera, "totalWidth", "Camera/totalWidth");
        if (!width)
            return false;
        if (almost_equal_relative_zero(*width))
            return fail("\"Camera/totalWidth\" should not be zero");
        const auto height = number(camera, "totalHeight", "Camera/totalHeight");
        if (!height)
            return false;
        if (almost_equal_relative_zero(*height))
            return fail("\"Camera/totalHeight\" should not be zero");
        const auto pw = number(camera, "pixelWidth", "Camera/pixelWidth");
        if (!pw)
            return false;
        if (*pw == 0.f)
            return fail("\"Camera/pixelWidth\" should not be zero");
        const auto ph = number(camera, "pixelHeight", "Camera/pixelHeight");
        if (!ph)
            return false;
        if (*ph == 0.f)
            return fail("\"Camera/pixelHeight\" should not be zero");
        int nu_full, nv_full;
        if (!to_int(*pw, nu_full) || !to_int(*ph, nv_full) || nu_full <= 0 || nv_full <= 0)
            return fail("detector pixel counts must be positive");
        // PixelSpacing::FromSizes (PixelSpacing.cpp:6-25)
        d.pixelSpacing = PixelSpacing(*width / (float)nu_full, *height / (float)nv_full);

        // per-view detector corners and heights (:307-339)
        const xml_lite::Node *refs = camera->first_child_element("references");
        if (!refs)
            return fail("no \"Camera/references\" element");
        std::vector<Position2D> corners;
        for (const xml_lite::Node *r : refs->child_elements()) {
            const std::string *t = r->get_text();
            const auto p = t ? parse_position(*t) : std::nullopt;
            if (!p)
                return fail("one among \"Camera/references\" elements could not be parsed");
            corners.emplace_back(p->x - center->x, p->y - center->y);
            d.detectorZs.push_back(p->z - center->z);
        }
        if (corners.empty())
            return fail("detectors positions empty");
        // the marshalling reads one source y per view (Projector.cu:927-932);
        // extra sources only enter the x and z means
        if (d.sources.size() < corners.size())
            return fail("fewer sources than detector references");

        // region read on every view (:353-430)
        const xml_lite::Node *radios = root->first_child_element("Radios");
        if (!radios)
            return fail("no \"Radios\" element");
        const xml_lite::Node *rois = radios->first_child_element("rois");
        if (!rois)
            return fail("no \"Radios/rois\" element");
        int bl[2] = {0, 0}, tr[2] = {0, 0};
        int index = 0;
        for (const xml_lite::Node *r : rois->child_elements()) {
            const std::string *t = r->get_text();
            const auto w = t ? words_of(*t) : std::vector<std::string>{};
            if (w.size() < 6)
                return fail("a \"Radios/rois\" entry needs six values");
            int v[4];
            const std::string *src[4] = {&w[0], &w[1], &w[4], &w[5]};
            for (int k = 0; k < 4; ++k) {
                const auto x = parse_double(*src[k]);
                if (!x || !to_int(*x, v[k]))
                    return fail("roi position parsing error");
            }
            if (index == 0) {
                bl[0] = v[0], bl[1] = v[1], tr[0] = v[2], tr[1] = v[3];
            } else if (bl[0] != v[0] || bl[1] != v[1]) {
                return fail("different roi bottom left positions");
            } else if (tr[0] != v[2] || tr[1] != v[3]) {
                return fail("different roi top right positions");
            }
            ++index;
        }
        bl[0] = std::clamp(bl[0], 0, nu_full - 1);
        bl[1] = std::clamp(bl[1], 0, nv_full - 1);
        tr[0] = std::clamp(tr[0], 0, nu_full - 1);
        tr[1] = std::clamp(tr[1], 0, nv_full - 1);
        if (tr[0] < bl[0] || tr[1] < bl[1])
            return fail("empty region of interest");
        d.roiSize = Size2D(tr[0] - bl[0] + 1, tr[1] - bl[1] + 1);
        d.roiPixel = Pixel(bl[0], bl[1]);
        for (const Position2D &c : corners) {
            // pixel positions are FP32 running sums of the pitch
            float x = c.x, y = c.y;
            for (int i = 0; i < bl[0]; ++i)
                x = x + d.pixelSpacing.x;
            for (int j = 0; j < bl[1]; ++j)
                y = y + d.pixelSpacing.y;
            d.roiBottomLefts.emplace_back(x, y);
        }
        d.ok = true;
        return true;
    }

    // AlmostEqualRelative(v, 0.F) of commons/Maths.h:25-34
    static bool almost_equal_relative_zero(float v)
    {
        const float diff = std::fabs(v - 0.f);
        const float a = std::fabs(v);
        return diff <= FLT_EPSILON * a;
    }
};

}  // namespace

TomoGeometryDescription ReadTomoGeometryXml(const std::string &xmlGeometryFilePath)
{
    Reader r;
    r.read(xmlGeometryFilePath);
    return r.d;
}

// Kernel arguments of an XML geometry file (the values Projector.cu:876-940
// passes to its launches): ints = {valid, nx, ny, nz, nu, nv, np}, floats =
// {sx, sy, sz, pu, pv, fulcrum x, y, z}, det/src = 3 floats per view for the
// first min(np, cap) views.  Returns np, or 0 with the reason in err when the
// file is not a valid geometry.
extern "C" int kvd_geometry_from_xml(const char *path, int32_t *ints, float *floats, float *det, float *src, int cap,
                                     char *err, int err_len)
{
    if (!path || !ints || !floats) {
        if (err && err_len > 0)
            std::snprintf(err, (size_t)err_len, "null argument");
        return 0;
    }
    const TomoGeometry tg{std::string(path)};
    std::memset(ints, 0, 7 * sizeof(int32_t));
    std::memset(floats, 0, 8 * sizeof(float));
    if (!tg.IsValid()) {
        if (err && err_len > 0)
            std::snprintf(err, (size_t)err_len, "%s", tg.ErrorMessage().c_str());
        return 0;
    }
    const MarshalledGeometry m = MarshalGeometry(&tg);
    const int32_t iv[7] = {1, m.g.nx, m.g.ny, m.g.nz, m.g.nu, m.g.nv, m.g.np};
    std::memcpy(ints, iv, sizeof iv);
    const WSize3D f = tg.fulcrum();
    const float fv[8] = {m.g.sx, m.g.sy, m.g.sz, m.g.pu, m.g.pv, f.x, f.y, f.z};
    std::memcpy(floats, fv, sizeof fv);
    const int n = std::min(m.g.np, std::max(cap, 0));
    if (det && n > 0)
        std::memcpy(det, m.det.data(), (size_t)n * 3 * sizeof(float));
    if (src && n > 0)
        std::memcpy(src, m.src.data(), (size_t)n * 3 * sizeof(float));
    if (err && err_len > 0)
        err[0] = '\0';
    return m.g.np;
}

extern "C" int kvd_geometry_roi_from_xml(const char *path, int32_t *roi, char *err, int err_len)
{
    if (!path || !roi) {
        if (err && err_len > 0)
            std::snprintf(err, (size_t)err_len, "null argument");
        return 1;
    }
    const TomoGeometry tg{std::string(path)};
    if (!tg.IsValid()) {
        if (err && err_len > 0)
            std::snprintf(err, (size_t)err_len, "%s", tg.ErrorMessage().c_str());
        return 1;
    }
    const Pixel bl = tg.GetProjectionROIsBLPixelPositionOnDetector();
    const Size2D size = tg.projectionsRoisSize();
    roi[0] = bl.x, roi[1] = bl.y, roi[2] = size.x, roi[3] = size.y;
    if (err && err_len > 0)
        err[0] = '\0';
    return 0;
}
