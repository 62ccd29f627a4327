// Drop-in Projector and recons:: over libkevernals_amd (see Reconstructors.h).
#include "Reconstructors.h"

#include <cstring>
#include <iostream>
#include <vector>

#include "Marshal.h"
#include "TiffWriter.h"
#include "kevernals_amd.h"

#ifdef KV_USE_REFERENCE_TREE
#ifdef KV_USE_REAL_VTK
// the reference's own PhantomMaker (its header needs VTK)
#include "modules/dataHandling/PhantomMaker.h"
#endif
#else
// ------------------------------------------------------------ error codes --
namespace {
class ReconstructorsCategory : public std::error_category {
public:
    const char *name() const noexcept override { return "Reconstructors"; }
    std::string message(int c) const override
    {
        switch ((ReconstructorsErrorCode)c) {
        case ReconstructorsErrorCode::ART: return "ART reconstruction failed";
        case ReconstructorsErrorCode::BackProjectionFailed: return "back projection failed";
        case ReconstructorsErrorCode::ProjectionFailed: return "projection failed";
        case ReconstructorsErrorCode::MLEM: return "MLEM reconstruction failed";
        case ReconstructorsErrorCode::ShiftAndAdd: return "shift and add reconstruction failed";
        }
        return "unknown reconstruction error";
    }
};
class ProjectorCategory : public std::error_category {
public:
    const char *name() const noexcept override { return "Projector"; }
    std::string message(int c) const override
    {
        switch ((ProjectorErrorCode)c) {
        case ProjectorErrorCode::InvalidGeometryError: return "invalid geometry";
        case ProjectorErrorCode::GeometryInconsistency: return "geometry inconsistency";
        case ProjectorErrorCode::VolumeProblemForProjection: return "volume problem for projection";
        case ProjectorErrorCode::ProjectionsDataFileReadingProblem: return "projections data reading problem";
        case ProjectorErrorCode::RadonMatrixComputation: return "radon matrix computation";
        case ProjectorErrorCode::BadAlgorithm: return "bad algorithm";
        }
        return "unknown projector error";
    }
};
const ReconstructorsCategory g_recons_category;
const ProjectorCategory g_projector_category;
}  // namespace

std::error_code make_error_code(ReconstructorsErrorCode e) { return {(int)e, g_recons_category}; }
std::error_code make_error_code(ProjectorErrorCode e) { return {(int)e, g_projector_category}; }
#endif  // KV_USE_REFERENCE_TREE

// ---------------------------------------------------------------- helpers --
namespace {
size_t points(vtkImageData *img)
{
    int d[3];
    img->GetDimensions(d);
    return (size_t)(d[0] > 0 ? d[0] : 0) * (size_t)(d[1] > 0 ? d[1] : 0) * (size_t)(d[2] > 0 ? d[2] : 0);
}

bool same_dims(vtkImageData *img, int x, int y, int z)
{
    int d[3];
    img->GetDimensions(d);
    return d[0] == x && d[1] == y && d[2] == z;
}

ImageDataPtr new_image(int x, int y, int z, double sx, double sy, double sz)
{
    auto img = ImageDataPtr::New();
    img->SetDimensions(x, y, z);
    img->SetSpacing(sx, sy, sz);
    img->AllocateScalars(VTK_FLOAT, 1);
    return img;
}

// Device buffer owned by a context.
struct DevBuf {
    kv_context *ctx = nullptr;
    void *p = nullptr;
    ~DevBuf()
    {
        if (p)
            kv_free(ctx, p);
    }
    bool alloc(kv_context *c, size_t bytes)
    {
        ctx = c;
        return kv_malloc(c, bytes, &p) == KV_OK;
    }
    float *f() const { return static_cast<float *>(p); }
};
}  // namespace

// -------------------------------------------------------------- Projector --
Projector::Projector(TomoGeometry *p_tomoGeometry) : m_tomoGeometry(p_tomoGeometry) {}

Projector::~Projector()
{
    if (ctx_)
        kv_destroy(ctx_);
}

bool Projector::ensureContext() const
{
    lastError_.clear();
    lastMessage_.clear();
    if (!m_tomoGeometry || !m_tomoGeometry->IsValid()) {
        lastError_ = make_error_code(ProjectorErrorCode::InvalidGeometryError);
        lastMessage_ = "invalid or missing geometry";
        return false;
    }
    if (ctx_)
        return true;
    kv_context *c = nullptr;
    if (kv_create(0, &c) != KV_OK) {
        lastError_ = make_error_code(ProjectorErrorCode::BadAlgorithm);
        lastMessage_ = kv_last_error();
        return false;
    }
    MarshalledGeometry m = MarshalGeometry(m_tomoGeometry);
    if (kv_set_geometry(c, &m.g) != KV_OK) {
        lastError_ = make_error_code(ProjectorErrorCode::InvalidGeometryError);
        lastMessage_ = kv_last_error();
        kv_destroy(c);
        return false;
    }
    ctx_ = c;
    return true;
}

kv_context *Projector::context() const { return ensureContext() ? ctx_ : nullptr; }

vtkSmartPointer<vtkImageData> Projector::PerformProjection(vtkSmartPointer<vtkImageData> p_volume) const
{
    if (!ensureContext())
        return nullptr;
    const Size3D vs = m_tomoGeometry->GetVolume()->GetSize3D();
    if (p_volume == nullptr || !p_volume->GetScalarPointer() || points(p_volume) != (size_t)vs.x * vs.y * vs.z) {
        lastError_ = make_error_code(ProjectorErrorCode::VolumeProblemForProjection);
        lastMessage_ = "volume image missing or of the wrong size";
        return nullptr;
    }
    const Size2D ds = m_tomoGeometry->projectionsRoisSize();
    const PixelSpacing ps = m_tomoGeometry->projectionsPixelSpacing();
    // Projector.cu:965-969
    auto out = new_image(ds.x, ds.y, m_tomoGeometry->nbProjectionsRois(), ps.x, ps.y, 1.);
    if (kv_forward_project_host(ctx_, static_cast<const float *>(p_volume->GetScalarPointer()),
                                static_cast<float *>(out->GetScalarPointer()), KV_MODE_R) != KV_OK) {
        lastError_ = make_error_code(ProjectorErrorCode::BadAlgorithm);
        lastMessage_ = kv_last_error();
        return nullptr;
    }
    return out;
}

void Projector::PerformProjection(const vtkSmartPointer<vtkImageData> p_volume, const Position3D &p_sourcePosition,
                                  vtkSmartPointer<vtkImageData> p_projectionsContainer) const
{
    lastError_.clear();
    lastMessage_.clear();
    if (!m_tomoGeometry || p_volume == nullptr || p_projectionsContainer == nullptr || !p_volume->GetScalarPointer() ||
        !p_projectionsContainer->GetScalarPointer()) {
        lastError_ = make_error_code(ProjectorErrorCode::VolumeProblemForProjection);
        lastMessage_ = "missing volume or container";
        return;
    }
    // Projector.cu:826-850: volume and detector sizes from the images
    int vd[3], pd[3];
    double vsp[3], psp[3];
    p_volume->GetDimensions(vd);
    p_volume->GetSpacing(vsp);
    p_projectionsContainer->GetDimensions(pd);
    p_projectionsContainer->GetSpacing(psp);
    const std::vector<Position2D> bl = m_tomoGeometry->projectionsRoisBottomLeftPositions();
    const int views = pd[2] > 0 ? pd[2] : 1;
    // Projector.cu:853-858: view p's detector origin is the geometry's region
    // corner of view p (.at(p): a container with more views than the
    // geometry is an error)
    if ((size_t)views > bl.size()) {
        lastError_ = make_error_code(ProjectorErrorCode::GeometryInconsistency);
        lastMessage_ = "container has more views than the geometry";
        return;
    }
    std::vector<float> det((size_t)views * 3), src((size_t)views * 3);
    for (int p = 0; p < views; ++p) {
        det[3 * p] = bl[(size_t)p].x;
        det[3 * p + 1] = bl[(size_t)p].y;
        det[3 * p + 2] = m_tomoGeometry->detectorsZCommonPosition();
        src[3 * p] = p_sourcePosition.x;
        src[3 * p + 1] = p_sourcePosition.y;
        src[3 * p + 2] = p_sourcePosition.z;
    }
    kv_geometry g{vd[0], vd[1], vd[2], (float)vsp[0], (float)vsp[1], (float)vsp[2], pd[0], pd[1], views,
                  (float)psp[0], (float)psp[1], det.data(), src.data()};
    kv_context *c = nullptr;
    if (kv_create(0, &c) != KV_OK || kv_set_geometry(c, &g) != KV_OK ||
        kv_forward_project_host(c, static_cast<const float *>(p_volume->GetScalarPointer()),
                                static_cast<float *>(p_projectionsContainer->GetScalarPointer()), KV_MODE_R) != KV_OK) {
        lastError_ = make_error_code(ProjectorErrorCode::BadAlgorithm);
        lastMessage_ = kv_last_error();
    }
    if (c)
        kv_destroy(c);
}

vtkSmartPointer<vtkImageData> Projector::PerformBackProjection(vtkSmartPointer<vtkImageData> p_projections) const
{
    if (!ensureContext())
        return nullptr;
    const Size2D ds = m_tomoGeometry->projectionsRoisSize();
    const size_t nray = (size_t)ds.x * ds.y * m_tomoGeometry->nbProjectionsRois();
    if (p_projections == nullptr || !p_projections->GetScalarPointer() || points(p_projections) != nray) {
        lastError_ = make_error_code(ProjectorErrorCode::GeometryInconsistency);
        lastMessage_ = "projection image missing or of the wrong size";
        return nullptr;
    }
    const Size3D vs = m_tomoGeometry->GetVolume()->GetSize3D();
    const VoxelSpacing sp = m_tomoGeometry->GetVolume()->GetVoxelSpacing();
    // Projector.cu:1136-1140
    auto out = new_image(vs.x, vs.y, vs.z, sp.x, sp.y, sp.z);
    if (kv_back_project_host(ctx_, static_cast<const float *>(p_projections->GetScalarPointer()),
                             static_cast<float *>(out->GetScalarPointer()), KV_MODE_R) != KV_OK) {
        lastError_ = make_error_code(ProjectorErrorCode::BadAlgorithm);
        lastMessage_ = kv_last_error();
        return nullptr;
    }
    return out;
}

#if defined(KV_USE_REFERENCE_TREE) && !defined(KV_USE_REAL_VTK)
namespace {
// Inside the reference tree without VTK the reference's PhantomMaker header
// cannot be included; this is its GetPhantom with no shapes
// (PhantomMaker.cpp:9-51): every voxel at the background density.
class PhantomMaker {
public:
    PhantomMaker(TomoVolume *p_volume, float p_backgroundDensity) : m_volume(p_volume), m_background(p_backgroundDensity)
    {
    }
    ImageDataPtr GetPhantom() const
    {
        if (!m_volume)
            return nullptr;
        const Size3D s = m_volume->GetSize3D();
        const VoxelSpacing sp = m_volume->GetVoxelSpacing();
        auto img = new_image(s.x, s.y, s.z, sp.x, sp.y, sp.z);
        float *d = static_cast<float *>(img->GetScalarPointer());
        std::fill(d, d + points(img), m_background);
        return img;
    }

private:
    TomoVolume *m_volume;
    float m_background;
};
}  // namespace
#endif

#ifndef KV_USE_REFERENCE_TREE
ImageDataPtr PhantomMaker::GetPhantom() const
{
    if (!m_volume)
        return nullptr;
    const Size3D s = m_volume->GetSize3D();
    const VoxelSpacing sp = m_volume->GetVoxelSpacing();
    auto img = new_image(s.x, s.y, s.z, sp.x, sp.y, sp.z);
    float *d = static_cast<float *>(img->GetScalarPointer());
    std::fill(d, d + points(img), m_backgroundDensity);
    return img;
}
#endif  // KV_USE_REFERENCE_TREE

// ----------------------------------------------------------------- recons --
namespace recons {
namespace {
enum class Loop { Art, Mlem };

// A verbose-mode dump: the image written as TIFF to the output directory
// path + name (the reference concatenates the two, Reconstructors.h:50).
void dump(const std::optional<std::string> &dir, const std::string &name, vtkImageData *img)
{
    if (!dir.has_value())
        return;
    auto writer = vtkSmartPointer<vtkTIFFWriter>::New();
    writer->SetFileName((dir.value() + name).c_str());
    writer->SetInputData(img);
    writer->Write();
    if (writer->GetErrorCode())
        std::cout << "TIFF dump " << dir.value() + name << " failed (error " << writer->GetErrorCode() << ")"
                  << std::endl;
}

// Iterations of recons::ART / recons::MLEM (Reconstructors.h:19-253).  With
// an output directory the steps run apart (forward projection, correction,
// backprojection, update) so that each intermediate image can be written
// (Reconstructors.h:47-53,76-82,96-102,125-131 and :164-170,193-199,
// 216-222,244-250); the values are those of the fused iterations.
Result<ImageDataPtr> iterate(Loop kind, TomoGeometry *tg, ImageDataPtr proj, int iterations, float relax,
                             std::optional<ImageDataPtr> initial, const std::optional<std::string> &outdir)
{
    // MLEM reports its failures with the ART code too (Reconstructors.h:159,190,230)
    const auto fail = ReconstructorsErrorCode::ART;
    ImageDataPtr vol;
    if (initial.has_value()) {
        vol = initial.value();  // mutated in place and returned (:31,120,240)
    } else {
        if (!tg)
            return fail;
        PhantomMaker maker(tg->GetVolume(), 1.f);  // :35-36, :152-153
        vol = maker.GetPhantom();
    }
    if (vol == nullptr || proj == nullptr || !tg)
        return fail;
    dump(outdir, "initialVolume.tiff", vol);
    Projector projector{tg};
    kv_context *ctx = projector.context();
    if (!ctx)
        return fail;
    const Size3D vs = tg->GetVolume()->GetSize3D();
    const Size2D ds = tg->projectionsRoisSize();
    const int np = tg->nbProjectionsRois();
    const size_t nvox = (size_t)vs.x * vs.y * vs.z, nray = (size_t)ds.x * ds.y * np;
    if (points(vol) != nvox || !vol->GetScalarPointer())
        return fail;
    float *x = static_cast<float *>(vol->GetScalarPointer());
    const bool data_ok = same_dims(proj, ds.x, ds.y, np) && proj->GetScalarPointer();
    const bool vol_ok = same_dims(vol, vs.x, vs.y, vs.z);

    if (data_ok && vol_ok) {
        // device-resident fused iterations
        DevBuf dv, dd;
        if (!dv.alloc(ctx, nvox * 4) || !dd.alloc(ctx, nray * 4))
            return fail;
        if (kv_memcpy_h2d(ctx, dv.p, x, nvox * 4) || kv_memcpy_h2d(ctx, dd.p, proj->GetScalarPointer(), nray * 4))
            return fail;
        if (!outdir.has_value()) {
            for (int it = 0; it < iterations; ++it) {
                const kv_status st = kind == Loop::Art ? kv_sart_iteration(ctx, dv.f(), dd.f(), relax)
                                                       : kv_mlem_iteration(ctx, dv.f(), dd.f(), relax, KV_MODE_R);
                if (st != KV_OK)
                    return fail;
            }
        } else {
            const int loop = kind == Loop::Art ? KV_LOOP_ART : KV_LOOP_MLEM;
            const PixelSpacing ps = tg->projectionsPixelSpacing();
            auto cur = new_image(ds.x, ds.y, np, ps.x, ps.y, 1.);  // Projector.cu:965-969
            float *c = static_cast<float *>(cur->GetScalarPointer());
            DevBuf dp, db;
            if (!dp.alloc(ctx, nray * 4) || !db.alloc(ctx, nvox * 4))
                return fail;
            for (int it = 0; it < iterations; ++it) {
                const std::string step = std::to_string(it) + ".tiff";
                if (kv_forward_project(ctx, dv.f(), dp.f(), KV_MODE_R) || kv_memcpy_d2h(ctx, c, dp.p, nray * 4))
                    return fail;
                dump(outdir, "currentProjectionStep" + step, cur);
                if (kv_correct_projections(ctx, loop, relax, dd.f(), dp.f()) || kv_memcpy_d2h(ctx, c, dp.p, nray * 4))
                    return fail;
                dump(outdir, "currentCorrectedProjectionStep" + step, cur);
                if (kv_back_project(ctx, dp.f(), db.f(), KV_MODE_R) || kv_apply_update(ctx, loop, dv.f(), db.f()) ||
                    kv_memcpy_d2h(ctx, x, dv.p, nvox * 4))
                    return fail;
                dump(outdir, "reconstructionstep" + step, vol);
            }
        }
        if (kv_memcpy_d2h(ctx, x, dv.p, nvox * 4))
            return fail;
        vol->Modified();
        return vol;
    }
    // A data image whose size differs from the projector output leaves the
    // projection unchanged before backprojection, and a volume of another
    // shape skips the update (Reconstructors.h:87,115,204,235).
    for (int it = 0; it < iterations; ++it) {
        const std::string step = std::to_string(it) + ".tiff";
        auto cur = projector.PerformProjection(vol);
        if (cur == nullptr)
            return fail;
        dump(outdir, "currentProjectionStep" + step, cur);
        float *c = static_cast<float *>(cur->GetScalarPointer());
        if (data_ok) {
            const float *p = static_cast<const float *>(proj->GetScalarPointer());
            for (size_t i = 0; i < nray; ++i) {
                if (kind == Loop::Art) {
                    c[i] = relax * (p[i] - c[i]);
                } else {
                    const float A = std::fabs(c[i]), diff = std::fabs(c[i] - 0.f);
                    if (!(diff <= 1.19209290e-07f * ((A < 0.f) ? 0.f : A)))  // !AlmostEqualRelative(c, 0)
                        c[i] = relax * (p[i] / c[i]);
                }
            }
        }
        dump(outdir, "currentCorrectedProjectionStep" + step, cur);
        auto back = projector.PerformBackProjection(cur);
        if (back == nullptr)
            return fail;
        if (vol_ok) {
            const float *b = static_cast<const float *>(back->GetScalarPointer());
            for (size_t i = 0; i < nvox; ++i)
                x[i] = kind == Loop::Art ? x[i] + b[i] : x[i] * b[i];
        }
        dump(outdir, "reconstructionstep" + step, vol);
    }
    return vol;
}
}  // namespace

Result<ImageDataPtr> ART(TomoGeometry *p_tomoGeometry, ImageDataPtr p_projectionImages, int p_iterationNumber,
                         float p_relaxationCoefficient, std::optional<ImageDataPtr> p_initialVolume,
                         std::optional<std::string> p_outputDirectoryPath)
{
    return iterate(Loop::Art, p_tomoGeometry, p_projectionImages, p_iterationNumber, p_relaxationCoefficient,
                   p_initialVolume, p_outputDirectoryPath);
}

Result<ImageDataPtr> MLEM(TomoGeometry *p_tomoGeometry, ImageDataPtr p_projectionImages, int p_iterationNumber,
                          float p_relaxationCoefficient, std::optional<ImageDataPtr> p_initialVolume,
                          std::optional<std::string> p_outputDirectoryPath)
{
    return iterate(Loop::Mlem, p_tomoGeometry, p_projectionImages, p_iterationNumber, p_relaxationCoefficient,
                   p_initialVolume, p_outputDirectoryPath);
}

Result<ImageDataPtr> BackProjection(TomoGeometry *p_tomoGeometry, ImageDataPtr p_projectionImages)
{
    Projector projector{p_tomoGeometry};
    auto out = projector.PerformBackProjection(p_projectionImages);
    if (out == nullptr)
        return ReconstructorsErrorCode::BackProjectionFailed;
    return out;
}

Result<ImageDataPtr> ShiftAndAdd(TomoGeometry *p_tomoGeometry, ImageDataPtr p_projectionImages,
                                 std::optional<std::string> p_outputDirectoryPath)
{
    if (!p_tomoGeometry || p_projectionImages == nullptr || !p_projectionImages->GetScalarPointer())
        return ReconstructorsErrorCode::ShiftAndAdd;
    if (p_outputDirectoryPath.has_value()) {  // Reconstructors.h:288-298
        std::cout << "zF = " << std::to_string(p_tomoGeometry->fulcrum().z) << std::endl;
        std::cout << "sid = " << std::to_string(p_tomoGeometry->sid()) << std::endl;
    }
    Projector projector{p_tomoGeometry};
    kv_context *ctx = projector.context();
    if (!ctx)
        return ReconstructorsErrorCode::ShiftAndAdd;
    const Size2D ds = p_tomoGeometry->projectionsRoisSize();
    const int np = p_tomoGeometry->nbProjections();
    if (!same_dims(p_projectionImages, ds.x, ds.y, np))
        return ReconstructorsErrorCode::ShiftAndAdd;
    const std::vector<float> zs = p_tomoGeometry->volumeZs();  // :317
    const std::vector<float> ys = p_tomoGeometry->sourcesYPositions();
    const size_t nray = (size_t)ds.x * ds.y * np, nout = (size_t)ds.x * ds.y * zs.size();
    double sp[3];
    p_projectionImages->GetSpacing(sp);
    auto out = new_image(ds.x, ds.y, (int)zs.size(), sp[0], sp[1], sp[2]);
    DevBuf dp, dout;
    if (!dp.alloc(ctx, nray * 4) || !dout.alloc(ctx, (nout ? nout : 1) * 4))
        return ReconstructorsErrorCode::ShiftAndAdd;
    if (kv_memcpy_h2d(ctx, dp.p, p_projectionImages->GetScalarPointer(), nray * 4))
        return ReconstructorsErrorCode::ShiftAndAdd;
    const kv_status st = kv_shift_and_add(ctx, dp.f(), ys.data(), (int)zs.size(), zs.data(), p_tomoGeometry->sid(),
                                          p_tomoGeometry->fulcrum().z, p_tomoGeometry->volumeVoxelSpacing().y,
                                          dout.f());
    if (st != KV_OK) {
        std::cout << "ShiftAndAdd: " << kv_last_error() << std::endl;
        return ReconstructorsErrorCode::ShiftAndAdd;
    }
    if (kv_memcpy_d2h(ctx, out->GetScalarPointer(), dout.p, nout * 4))
        return ReconstructorsErrorCode::ShiftAndAdd;
    return out;
}
Result<ImageDataPtr> MLEMAccelerated(TomoGeometry *p_tomoGeometry, ImageDataPtr p_lineIntegrals,
                                     std::optional<ImageDataPtr> p_initialVolume,
                                     const std::vector<float> *p_cnnParameters, float p_cnnScale,
                                     AcceleratedSchedule p_schedule, int *p_stepsKept)
{
    const auto fail = ReconstructorsErrorCode::MLEM;
    if (!p_tomoGeometry || p_lineIntegrals == nullptr)
        return fail;
    ImageDataPtr vol;
    if (p_initialVolume.has_value()) {
        vol = p_initialVolume.value();
    } else {
        PhantomMaker maker(p_tomoGeometry->GetVolume(), 1.f);
        vol = maker.GetPhantom();
    }
    if (vol == nullptr || !vol->GetScalarPointer() || !p_lineIntegrals->GetScalarPointer())
        return fail;
    const Size3D vs = p_tomoGeometry->GetVolume()->GetSize3D();
    const Size2D ds = p_tomoGeometry->projectionsRoisSize();
    const int np = p_tomoGeometry->nbProjectionsRois();
    if (!same_dims(vol, vs.x, vs.y, vs.z) || !same_dims(p_lineIntegrals, ds.x, ds.y, np))
        return fail;
    if (p_cnnParameters && p_cnnParameters->size() != (size_t)KV_CNN_NPARAMS)
        return fail;
    Projector projector{p_tomoGeometry};
    kv_context *ctx = projector.context();
    if (!ctx)
        return fail;
    const size_t nvox = (size_t)vs.x * vs.y * vs.z, nray = (size_t)ds.x * ds.y * np;
    DevBuf dv, dd, dw;
    if (!dv.alloc(ctx, nvox * 4) || !dd.alloc(ctx, nray * 4))
        return fail;
    if (kv_memcpy_h2d(ctx, dv.p, vol->GetScalarPointer(), nvox * 4) ||
        kv_memcpy_h2d(ctx, dd.p, p_lineIntegrals->GetScalarPointer(), nray * 4))
        return fail;
    if (p_cnnParameters) {
        if (!dw.alloc(ctx, (size_t)KV_CNN_NPARAMS * 4) ||
            kv_memcpy_h2d(ctx, dw.p, p_cnnParameters->data(), (size_t)KV_CNN_NPARAMS * 4))
            return fail;
    }
    const kv_accel_schedule sch{p_schedule.start, p_schedule.first, p_schedule.mid, p_schedule.second,
                                p_schedule.horizon};
    int kept = 0;
    const kv_status st =
        kv_mlem_accelerated(ctx, dv.f(), dd.f(), &sch, p_cnnParameters ? dw.f() : nullptr, p_cnnScale, &kept);
    if (st != KV_OK) {
        std::cout << "MLEMAccelerated: " << kv_last_error() << std::endl;
        return fail;
    }
    if (kv_memcpy_d2h(ctx, vol->GetScalarPointer(), dv.p, nvox * 4))
        return fail;
    vol->Modified();
    if (p_stepsKept)
        *p_stepsKept = kept;
    return vol;
}
}  // namespace recons
