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
