// The reference's own geometry reading, as a checker: TEST INFRASTRUCTURE
// ONLY.  Built by oracle/Makefile into oracle/_ref/ from the reference's
// sources where they lie (source/modules/geometry/*.cpp and the vendored
// commons/tinyXML/tinyxml2.cpp); only tests/ load it.
//
// ref_geometry_from_xml reads an XML geometry file with the reference's
// TomoGeometry (TomoGeometry.cpp:13-437) and gathers the kernel arguments as
// Projector::PerformProjection does (Projector.cu:876-936): ints = {valid,
// nx, ny, nz, nu, nv, np}, floats = {sx, sy, sz, pu, pv, fulcrum x, y, z},
// det/src = 3 floats per view for the first min(np, cap) views.  Returns np,
// or 0 for a file the reference does not accept.  The reference prints its
// progress on std::cout; that output is discarded here.
#include <algorithm>
#include <cstdint>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "modules/geometry/TomoGeometry.h"

extern "C" int ref_geometry_from_xml(const char *path, int32_t *ints, float *floats, float *det, float *src, int cap)
{
    std::ostringstream sink;
    std::streambuf *old = std::cout.rdbuf(sink.rdbuf());
    TomoGeometry tg{std::string(path)};
    std::cout.rdbuf(old);
    for (int i = 0; i < 7; ++i)
        ints[i] = 0;
    for (int i = 0; i < 8; ++i)
        floats[i] = 0.f;
    if (!tg.IsValid())
        return 0;
    const int np = tg.nbProjectionsRois();
    const auto vs = tg.GetVolume()->GetSize3D();
    const auto sp = tg.GetVolume()->GetVoxelSpacing();
    const auto ds = tg.projectionsRoisSize();
    const auto ps = tg.projectionsPixelSpacing();
    const auto bl = tg.projectionsRoisBottomLeftPositions();
    const std::vector<float> sy = tg.sourcesYPositions();
    const auto fc = tg.fulcrum();
    const int32_t iv[7] = {1, vs.x, vs.y, vs.z, ds.x, ds.y, np};
    const float fv[8] = {sp.x, sp.y, sp.z, ps.x, ps.y, fc.x, fc.y, fc.z};
    std::copy(iv, iv + 7, ints);
    std::copy(fv, fv + 8, floats);
    const int n = std::min(np, std::max(cap, 0));
    for (int p = 0; p < n; ++p) {
        det[3 * p + 0] = bl.at((size_t)p).x;
        det[3 * p + 1] = bl.at((size_t)p).y;
        det[3 * p + 2] = tg.detectorsZCommonPosition();
        src[3 * p + 0] = tg.sourcesXCommonPosition();
        src[3 * p + 1] = sy.at((size_t)p);
        src[3 * p + 2] = tg.sourcesZCommonPosition();
    }
    return np;
}

#include "modules/geometry/Pave.h"
#include "modules/geometry/Sphere.h"
#include "modules/geometry/TomoVolume.h"

// PhantomMaker::GetPhantom's densities (PhantomMaker.cpp:13-29) on the
// reference's own TomoVolume grid, Pave::Contains and Sphere::Contains: the
// volume is built as TomoGeometry.cpp:121-127 builds it from a size and a
// spacing; paves are 7 floats (corner, size, density), spheres 5 (centre,
// radius, density).  PhantomMaker.cpp itself needs VTK for its output image;
// its loop is restated here around the reference's geometry classes, which
// decide every bit.  Returns the voxel count.
extern "C" int ref_phantom(int nx, int ny, int nz, float sx, float sy, float sz, float background,
                           const float *paves, int npaves, const float *spheres, int nspheres, float *out)
{
    const VoxelSpacing spacing(sx, sy, sz);
    const Size3D size(nx, ny, nz);
    const WSize3D wsize(spacing, size);
    const Position3D blf(-0.5 * wsize.x, -0.5 * wsize.y, -0.5 * wsize.z);
    TomoVolume volume(blf, size, spacing);
    const auto grid = volume.GetGridPositions();
    std::vector<Pave> pv;
    std::vector<Sphere> sp;
    for (int k = 0; k < npaves; ++k) {
        const float *q = paves + 7 * k;
        pv.emplace_back(Position3D(q[0], q[1], q[2]), WSize3D(q[3], q[4], q[5]));
    }
    for (int k = 0; k < nspheres; ++k) {
        const float *q = spheres + 5 * k;
        sp.emplace_back(Position3D(q[0], q[1], q[2]), q[3]);
    }
    for (size_t i = 0; i < grid.size(); ++i) {
        float d = background;
        for (int k = 0; k < npaves; ++k)
            if (pv[(size_t)k].Contains(grid[i]))
                d += paves[7 * k + 6];
        for (int k = 0; k < nspheres; ++k)
            if (sp[(size_t)k].Contains(grid[i]))
                d += spheres[5 * k + 4];
        out[i] = d;
    }
    return (int)grid.size();
}
