// Runs the drop-in Projector and recons:: the way the reference's callers do
// (mainReconstruct.cpp:170-219) on inputs read from a directory, and writes
// the results as raw float32 files for comparison.
//
//   kv_dropin_check <dir>
//   <dir>/geom.xml : the reference's XML geometry file (TomoGeometry(path),
//                    TomoGeometry.cpp:13-437), or else
//   <dir>/geom.txt : nx ny nz sx sy sz nu nv np pu pv fulcrum_z, then per view
//                    src_x src_y src_z det_x det_y det_z
//   <dir>/vol.bin, <dir>/proj.bin : float32 volume and projections
//   <dir>/ell.bin (optional): float32 line integrals for the accelerated
//                    MLEM; <dir>/cnn.bin (optional): its CNN parameters
// Writes fp.bin, bp.bin, mlem.bin, mlem_default.bin, art.bin, saa.bin,
// mlem_verbose.bin and art_verbose.bin (with TIFF dumps under dump_mlem/ and
// dump_art/), accel.bin and accel_cnn.bin (with ell.bin / cnn.bin) and
// status.txt (one "name code" line per call).
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "Reconstructors.h"
#include "kevernals_amd.h"

namespace {
bool read_raw(const std::string &path, float *dst, size_t n)
{
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f)
        return false;
    const size_t got = std::fread(dst, sizeof(float), n, f);
    std::fclose(f);
    return got == n;
}

void write_raw(const std::string &path, ImageDataPtr img)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f)
        return;
    int d[3];
    img->GetDimensions(d);
    std::fwrite(img->GetScalarPointer(), sizeof(float), (size_t)d[0] * d[1] * d[2], f);
    std::fclose(f);
}

ImageDataPtr image(int x, int y, int z, double sx, double sy, double sz)
{
    auto img = ImageDataPtr::New();
    img->SetDimensions(x, y, z);
    img->SetSpacing(sx, sy, sz);
    img->AllocateScalars(VTK_FLOAT, 1);
    return img;
}
}  // namespace

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::cerr << "usage: kv_dropin_check <dir>" << std::endl;
        return 2;
    }
    const std::string dir = argv[1];
    std::vector<Position3D> src;
    std::vector<Position2D> bl;
    std::vector<float> dz;
    std::unique_ptr<TomoGeometry> gp;
    const bool from_xml = std::ifstream(dir + "/geom.xml").good();
    if (from_xml) {
        gp = std::make_unique<TomoGeometry>(dir + "/geom.xml");
        if (!gp->IsValid()) {
            std::cerr << "bad geom.xml: " << gp->ErrorMessage() << std::endl;
            return 2;
        }
    } else {
        std::ifstream in(dir + "/geom.txt");
        int tnx, tny, tnz, tnu, tnv, tnp;
        float tsx, tsy, tsz, tpu, tpv, tfz;
        if (!(in >> tnx >> tny >> tnz >> tsx >> tsy >> tsz >> tnu >> tnv >> tnp >> tpu >> tpv >> tfz)) {
            std::cerr << "bad geom.txt" << std::endl;
            return 2;
        }
        src.resize((size_t)tnp);
        bl.resize((size_t)tnp);
        dz.resize((size_t)tnp);
        for (int p = 0; p < tnp; ++p) {
            float a, b, c, d, e, f;
            in >> a >> b >> c >> d >> e >> f;
            src[(size_t)p] = Position3D(a, b, c);
            bl[(size_t)p] = Position2D(d, e);
            dz[(size_t)p] = f;
        }
        gp = std::make_unique<TomoGeometry>(Size3D(tnx, tny, tnz), VoxelSpacing(tsx, tsy, tsz), Size2D(tnu, tnv),
                                            PixelSpacing(tpu, tpv), src, bl, dz, WSize3D(0.f, 0.f, tfz));
    }
    TomoGeometry &geom = *gp;
    const Size3D vs = geom.volumeSize3D();
    const VoxelSpacing vsp = geom.volumeVoxelSpacing();
    const Size2D ds = geom.projectionsRoisSize();
    const PixelSpacing psp = geom.projectionsPixelSpacing();
    const int nx = vs.x, ny = vs.y, nz = vs.z, nu = ds.x, nv = ds.y, np = geom.nbProjectionsRois();
    const float sx = vsp.x, sy = vsp.y, sz = vsp.z, pu = psp.x, pv = psp.y;
    auto vol = image(nx, ny, nz, sx, sy, sz);
    auto proj = image(nu, nv, np, pu, pv, 1.0);
    if (!read_raw(dir + "/vol.bin", static_cast<float *>(vol->GetScalarPointer()), (size_t)nx * ny * nz) ||
        !read_raw(dir + "/proj.bin", static_cast<float *>(proj->GetScalarPointer()), (size_t)nu * nv * np)) {
        std::cerr << "cannot read inputs" << std::endl;
        return 2;
    }
    std::ofstream status(dir + "/status.txt");
    int failures = 0;
    auto note = [&](const char *name, std::error_code ec) {
        status << name << " " << ec.value() << std::endl;
        if (ec) {
            ++failures;
            std::cerr << name << ": " << ec.message() << std::endl;
        }
    };

    Projector projector(&geom);
    auto fp = projector.PerformProjection(vol);
    note("projection", fp == nullptr ? projector.lastError() : std::error_code());
    if (fp != nullptr)
        write_raw(dir + "/fp.bin", fp);
    auto bp = projector.PerformBackProjection(proj);
    note("backprojection", bp == nullptr ? projector.lastError() : std::error_code());
    if (bp != nullptr)
        write_raw(dir + "/bp.bin", bp);

    // copies of the volume: ART/MLEM mutate their initial volume (Q21)
    auto copy = [&](ImageDataPtr src_img) {
        auto c = image(nx, ny, nz, sx, sy, sz);
        std::copy((float *)src_img->GetScalarPointer(), (float *)src_img->GetScalarPointer() + (size_t)nx * ny * nz,
                  (float *)c->GetScalarPointer());
        return c;
    };
    auto mlem = recons::MLEM(&geom, proj, 3, 1.0f, copy(vol), std::nullopt);
    note("mlem", mlem.has_value() ? std::error_code() : mlem.error());
    if (mlem)
        write_raw(dir + "/mlem.bin", mlem.value());
    auto mlem0 = recons::MLEM(&geom, proj, 2, 1.0f, std::nullopt, std::nullopt);
    note("mlem_default", mlem0.has_value() ? std::error_code() : mlem0.error());
    if (mlem0)
        write_raw(dir + "/mlem_default.bin", mlem0.value());
    auto art = recons::ART(&geom, proj, 2, 0.5f, copy(vol), std::nullopt);
    note("art", art.has_value() ? std::error_code() : art.error());
    if (art)
        write_raw(dir + "/art.bin", art.value());
    // verbose mode: the TIFF dumps of every step (Reconstructors.h:47-131,164-250)
    {
        const std::string dm = dir + "/dump_mlem/", da = dir + "/dump_art/";
        std::filesystem::create_directories(dm);
        std::filesystem::create_directories(da);
        auto vm = recons::MLEM(&geom, proj, 3, 1.0f, copy(vol), dm);
        note("mlem_verbose", vm.has_value() ? std::error_code() : vm.error());
        if (vm)
            write_raw(dir + "/mlem_verbose.bin", vm.value());
        auto va = recons::ART(&geom, proj, 2, 0.5f, copy(vol), da);
        note("art_verbose", va.has_value() ? std::error_code() : va.error());
        if (va)
            write_raw(dir + "/art_verbose.bin", va.value());
    }
    auto bpr = recons::BackProjection(&geom, proj);
    note("recons_backprojection", bpr.has_value() ? std::error_code() : bpr.error());
    auto saa = recons::ShiftAndAdd(&geom, proj, std::nullopt);
    note("saa", saa.has_value() ? std::error_code() : saa.error());
    if (saa)
        write_raw(dir + "/saa.bin", saa.value());

    // single-source projection (Projector.cu:823-870) from view 0's
    // marshalled source position onto a container of every view
    {
        auto cont = image(nu, nv, np, pu, pv, 1.0);
        const Position3D s0(geom.sourcesXCommonPosition(), geom.sourcesYPositions().at(0),
                            geom.sourcesZCommonPosition());
        projector.PerformProjection(vol, s0, cont);
        note("single_source", projector.lastError());
        if (!projector.lastError())
            write_raw(dir + "/single.bin", cont);
    }

    // accelerated MLEM on line integrals (<dir>/ell.bin), without and, given
    // <dir>/cnn.bin (KV_CNN_NPARAMS floats), with the CNN step
    {
        auto ell = image(nu, nv, np, pu, pv, 1.0);
        if (read_raw(dir + "/ell.bin", static_cast<float *>(ell->GetScalarPointer()), (size_t)nu * nv * np)) {
            int kept = -1;
            auto acc = recons::MLEMAccelerated(&geom, ell, std::nullopt, nullptr, 20.f, {}, &kept);
            note("mlem_accelerated", acc.has_value() ? std::error_code() : acc.error());
            if (acc)
                write_raw(dir + "/accel.bin", acc.value());
            status << "accel_kept " << kept << std::endl;
            std::vector<float> params(KV_CNN_NPARAMS);
            if (read_raw(dir + "/cnn.bin", params.data(), params.size())) {
                kept = -1;
                auto accc = recons::MLEMAccelerated(&geom, ell, std::nullopt, &params, 20.f, {}, &kept);
                note("mlem_accelerated_cnn", accc.has_value() ? std::error_code() : accc.error());
                if (accc)
                    write_raw(dir + "/accel_cnn.bin", accc.value());
                status << "accel_cnn_kept " << kept << std::endl;
            }
        }
    }

    // error path: one reconstructed slice (Reconstructors.h:274-279)
    if (!from_xml) {
        TomoGeometry flat(Size3D(nx, ny, 1), VoxelSpacing(sx, sy, sz), Size2D(nu, nv), PixelSpacing(pu, pv), src, bl,
                          dz);
        auto bad = recons::ShiftAndAdd(&flat, proj, std::nullopt);
        status << "saa_one_slice " << (bad.has_value() ? 0 : bad.error().value()) << std::endl;
    }
    return failures == 0 ? 0 : 1;
}
