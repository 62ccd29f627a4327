// Acquisition geometry as the projector and reconstruction loops read it.
//
// The accessors mirror KevLoq/Kevernals source/modules/geometry/TomoGeometry.h
// (:34-86) and the values they return follow the reference's construction:
//  * the volume is centred on the origin, bottom-left-front = -0.5 * size
//    (TomoGeometry.cpp:125-127);
//  * source x and z are averaged over the views in FP32, sequentially
//    (TomoTable.cpp:11-18), while each view keeps its own source y;
//  * the detector plane height is the mean of the per-view heights
//    (TomoGeometry.cpp:345);
//  * voxel and pixel position tables are sequential FP32 sums
//    (TomoVolume.cpp:197-215, TomoProjectionsSet.cpp:183-194).
// A geometry is built from values, or read from the reference's XML
// geometry file (TomoGeometry.cpp:13-437) by TomoGeometry(path), whose
// reading rules are in TomoGeometryXml.cpp.
#pragma once

#include <cmath>
#include <cstddef>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

struct Position2D {
    float x = 0.f, y = 0.f;
    Position2D() = default;
    Position2D(float x_, float y_) : x(x_), y(y_) {}
};
struct Position3D {
    float x = 0.f, y = 0.f, z = 0.f;
    Position3D() = default;
    Position3D(float x_, float y_, float z_) : x(x_), y(y_), z(z_) {}
};
using WSize3D = Position3D;
// A detector pixel index (geometry/Pixel.h).
struct Pixel {
    int x = 0, y = 0;
    Pixel() = default;
    Pixel(int x_, int y_) : x(x_), y(y_) {}
};
struct Size2D {
    int x = 0, y = 0;
    Size2D() = default;
    Size2D(int x_, int y_) : x(x_), y(y_) {}
};
struct Size3D {
    int x = 0, y = 0, z = 0;
    Size3D() = default;
    Size3D(int x_, int y_, int z_) : x(x_), y(y_), z(z_) {}
};
struct PixelSpacing {
    float x = 1.f, y = 1.f;
    PixelSpacing() = default;
    PixelSpacing(float x_, float y_) : x(x_), y(y_) {}
};
struct VoxelSpacing {
    float x = 1.f, y = 1.f, z = 1.f;
    VoxelSpacing() = default;
    VoxelSpacing(float x_, float y_, float z_) : x(x_), y(y_), z(z_) {}
};

class TomoVolume {
public:
    TomoVolume(const Size3D &size, const VoxelSpacing &spacing) : size_(size), spacing_(spacing)
    {
        // WSize3D(voxelSpacing, resolution); BLF = -0.5 * wsize
        const float wx = spacing.x * (float)size.x, wy = spacing.y * (float)size.y, wz = spacing.z * (float)size.z;
        blf_ = Position3D((float)(-0.5 * wx), (float)(-0.5 * wy), (float)(-0.5 * wz));
    }
    Size3D GetSize3D() const { return size_; }
    VoxelSpacing GetVoxelSpacing() const { return spacing_; }
    Position3D GetBottomLeftFront() const { return blf_; }
    int GetVoxelsNumber() const { return size_.x * size_.y * size_.z; }
    // lower-face positions, sequential FP32 sums from the bottom-left-front
    std::vector<float> GetXPositions() const { return positions(blf_.x, spacing_.x, size_.x); }
    std::vector<float> GetYPositions() const { return positions(blf_.y, spacing_.y, size_.y); }
    std::vector<float> GetZPositions() const { return positions(blf_.z, spacing_.z, size_.z); }

private:
    static std::vector<float> positions(float start, float step, int n)
    {
        std::vector<float> p((size_t)(n > 0 ? n : 0));
        if (n > 0)
            p[0] = start;
        for (int i = 1; i < n; ++i)
            p[(size_t)i] = p[(size_t)i - 1] + step;
        return p;
    }
    Size3D size_;
    VoxelSpacing spacing_;
    Position3D blf_;
};

// What an XML geometry file describes, in the form the value constructor of
// TomoGeometry takes (TomoGeometryXml.cpp).
struct TomoGeometryDescription;

class TomoGeometry {
public:
    // Reads the reference's XML geometry file (TomoGeometry.cpp:13-437);
    // IsValid() is false and ErrorMessage() says why when the file is not a
    // valid geometry.
    explicit TomoGeometry(const std::string &xmlGeometryFilePath);

    // sources: per-view source positions (world mm, volume-centred frame;
    // at least one per view, extra ones only enter the x and z means);
    // detectorBottomLefts: per-view bottom-left corner of the detector region
    // read (the ROI); detectorZs: per-view detector plane heights; fulcrum:
    // the reference's Rotation/center (used by shift-and-add).
    TomoGeometry(const Size3D &volumeSize, const VoxelSpacing &voxelSpacing, const Size2D &detectorSize,
                 const PixelSpacing &pixelSpacing, const std::vector<Position3D> &sources,
                 const std::vector<Position2D> &detectorBottomLefts, const std::vector<float> &detectorZs,
                 const WSize3D &fulcrum = WSize3D())
        : volume_(std::make_unique<TomoVolume>(volumeSize, voxelSpacing)), detSize_(detectorSize),
          pixelSpacing_(pixelSpacing), roiBL_(detectorBottomLefts), fulcrum_(fulcrum)
    {
        // TomoTable.cpp:4-19
        sourcesY_.resize(sources.size());
        float sx = 0.f, sz = 0.f;
        for (size_t i = 0; i < sources.size(); ++i) {
            sx += sources[i].x;
            sourcesY_[i] = sources[i].y;
            sz += sources[i].z;
        }
        srcX_ = sources.empty() ? 0.f : sx / (float)sources.size();
        srcZ_ = sources.empty() ? 0.f : sz / (float)sources.size();
        // TomoGeometry.cpp:345
        detZ_ = detectorZs.empty()
                    ? 0.f
                    : (float)std::accumulate(detectorZs.begin(), detectorZs.end(), .0) / (float)detectorZs.size();
        valid_ = !detectorBottomLefts.empty() && sources.size() >= detectorBottomLefts.size() && volumeSize.x > 0 &&
                 volumeSize.y > 0 && volumeSize.z > 0 && detectorSize.x > 0 && detectorSize.y > 0;
    }

    bool IsValid() const { return valid_; }
    const std::string &ErrorMessage() const { return error_; }
    std::string filePath() const { return filePath_; }
    TomoVolume *GetVolume() const { return volume_.get(); }

    // sources
    std::vector<float> sourcesYPositions() const { return sourcesY_; }
    float sourcesXCommonPosition() const { return srcX_; }
    float sourcesZCommonPosition() const { return srcZ_; }
    float detectorsZCommonPosition() const { return detZ_; }
    float sid() const { return std::fabs(detZ_ - srcZ_); }

    // detector region read by the projector (the ROI of the reference)
    int nbProjections() const { return (int)roiBL_.size(); }
    int nbProjectionsRois() const { return (int)roiBL_.size(); }
    Size2D projectionsRoisSize() const { return detSize_; }
    PixelSpacing projectionsPixelSpacing() const { return pixelSpacing_; }
    std::vector<Position2D> projectionsRoisBottomLeftPositions() const { return roiBL_; }
    // bottom-left pixel of the region on the full detector (TomoGeometry.cpp:426)
    Pixel GetProjectionROIsBLPixelPositionOnDetector() const { return roiPixel_; }

    // volume
    Size3D volumeSize3D() const { return volume_->GetSize3D(); }
    VoxelSpacing volumeVoxelSpacing() const { return volume_->GetVoxelSpacing(); }
    Position3D volumeBLF() const { return volume_->GetBottomLeftFront(); }
    std::vector<float> volumeZs() const { return volume_->GetZPositions(); }
    WSize3D fulcrum() const { return fulcrum_; }

private:
    std::unique_ptr<TomoVolume> volume_;
    Size2D detSize_;
    PixelSpacing pixelSpacing_;
    std::vector<Position2D> roiBL_;
    Pixel roiPixel_;
    std::vector<float> sourcesY_;
    float srcX_ = 0.f, srcZ_ = 0.f, detZ_ = 0.f;
    WSize3D fulcrum_;
    bool valid_ = false;
    std::string error_, filePath_;

    explicit TomoGeometry(const TomoGeometryDescription &d);
};

struct TomoGeometryDescription {
    bool ok = false;
    std::string error;
    Size3D volumeSize;
    VoxelSpacing voxelSpacing;
    Size2D roiSize;
    Pixel roiPixel;
    PixelSpacing pixelSpacing;
    std::vector<Position3D> sources;
    std::vector<Position2D> roiBottomLefts;
    std::vector<float> detectorZs;
    WSize3D fulcrum;
};

// Reading rules of the reference's XML geometry file (TomoGeometryXml.cpp).
TomoGeometryDescription ReadTomoGeometryXml(const std::string &xmlGeometryFilePath);

inline TomoGeometry::TomoGeometry(const TomoGeometryDescription &d)
    : TomoGeometry(d.volumeSize, d.voxelSpacing, d.roiSize, d.pixelSpacing, d.sources, d.roiBottomLefts,
                   d.detectorZs, d.fulcrum)
{
    valid_ = valid_ && d.ok;
    roiPixel_ = d.roiPixel;
    error_ = d.ok ? (valid_ ? std::string() : std::string("geometry is empty")) : d.error;
}

inline TomoGeometry::TomoGeometry(const std::string &xmlGeometryFilePath)
    : TomoGeometry(ReadTomoGeometryXml(xmlGeometryFilePath))
{
    filePath_ = xmlGeometryFilePath;
}
