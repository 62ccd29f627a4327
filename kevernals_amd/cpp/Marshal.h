// Kernel arguments from a geometry's accessors, as Projector.cu:876-940
// gathers them before its launches: volume size and spacing, the ROI size
// and pixel pitch, per view the ROI's bottom-left corner at the common
// detector height and the source at the common x and z with the view's y.
// Works with any geometry type that has the reference's accessor names.
#pragma once

#include <vector>

#include "kevernals_amd.h"

struct MarshalledGeometry {
    kv_geometry g{};
    std::vector<float> det, src;  // g.det / g.src point into these
};

template <class Geometry>
MarshalledGeometry MarshalGeometry(const Geometry *tg)
{
    MarshalledGeometry m;
    const auto vs = tg->GetVolume()->GetSize3D();
    const auto sp = tg->GetVolume()->GetVoxelSpacing();
    const int np = tg->nbProjectionsRois();
    const auto ds = tg->projectionsRoisSize();
    const auto ps = tg->projectionsPixelSpacing();
    const auto bl = tg->projectionsRoisBottomLeftPositions();
    const std::vector<float> sy = tg->sourcesYPositions();
    m.det.resize((size_t)np * 3);
    m.src.resize((size_t)np * 3);
    for (int p = 0; p < np; ++p) {
        m.det[3 * p + 0] = bl.at((size_t)p).x;
        m.det[3 * p + 1] = bl.at((size_t)p).y;
        m.det[3 * p + 2] = tg->detectorsZCommonPosition();
        m.src[3 * p + 0] = tg->sourcesXCommonPosition();
        m.src[3 * p + 1] = sy.at((size_t)p);
        m.src[3 * p + 2] = tg->sourcesZCommonPosition();
    }
    m.g.nx = vs.x;
    m.g.ny = vs.y;
    m.g.nz = vs.z;
    m.g.sx = sp.x;
    m.g.sy = sp.y;
    m.g.sz = sp.z;
    m.g.nu = ds.x;
    m.g.nv = ds.y;
    m.g.np = np;
    m.g.pu = ps.x;
    m.g.pv = ps.y;
    m.g.det = m.det.data();
    m.g.src = m.src.data();
    return m;
}
