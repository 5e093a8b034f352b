// Checkpoint/restart (the reference has none): a binary file holding the
// hierarchy of every species, f slots 0 and 1 of every patch (whole planes,
// ghost rings included), the field state and the simulation clock.  A run
// restored from it continues bit for bit as the run that wrote it.
#include <cstring>
#include <fstream>

#include "veritas.hpp"

namespace vxh {

namespace {
const char kMagic[8] = {'V', 'X', 'C', 'K', 'P', 'T', '0', '1'};

struct Writer {
    std::ofstream o;
    void i32(int32_t v) { o.write((const char *)&v, sizeof v); }
    void f64(double v) { o.write((const char *)&v, sizeof v); }
    void vec(const std::vector<double> &v) {
        i32((int32_t)v.size());
        o.write((const char *)v.data(), sizeof(double) * v.size());
    }
};
struct Reader {
    std::ifstream in;
    int32_t i32() {
        int32_t v = 0;
        in.read((char *)&v, sizeof v);
        if (!in) throw Error("checkpoint: truncated file");
        return v;
    }
    double f64() {
        double v = 0;
        in.read((char *)&v, sizeof v);
        if (!in) throw Error("checkpoint: truncated file");
        return v;
    }
    std::vector<double> vec(size_t expect) {
        int32_t n = i32();
        if ((size_t)n != expect) throw Error("checkpoint: array size does not match this run");
        std::vector<double> v(n);
        in.read((char *)v.data(), sizeof(double) * n);
        if (!in) throw Error("checkpoint: truncated file");
        return v;
    }
};
}  // namespace

void Rectangle::UploadSlot(int slot, const std::vector<double> &full) {
    UnaliasSlots();
    if (full.size() != vx_plane_elems(n_x, n_p)) throw Error("checkpoint: plane size mismatch");
    for (double v : full) fScale = std::max(fScale, std::fabs(v));
    dev->check(vx_memcpy_h2d(dev->ctx, P.f[slot], full.data(), sizeof(double) * full.size()), "upload slot");
}

void Rectangle::UploadSlot(int slot, int g, const double *host) {
    UnaliasSlots();
    const size_t n = (size_t)(n_x + 2 * g) * (n_p + 2 * g);
    for (size_t k = 0; k < n; k++) fScale = std::max(fScale, std::fabs(host[k]));
    dev->check(vx_patch_upload(dev->ctx, &P, slot, g, host), "upload slot");
}

void Mesh::Restore(const std::vector<level> &h) {
    hierarchy = h;
    promoteHierarchyToMesh(true);
}

void SolverManager::SaveCheckpoint(const std::string &path) {
    Writer w;
    w.o.open(path, std::ios::binary);
    if (!w.o) throw Error("checkpoint: cannot open " + path);
    w.o.write(kMagic, sizeof kMagic);
    w.i32((int32_t)settings.x_size);
    w.i32((int32_t)settings.maxDepth);
    w.i32((int32_t)settings.refinementRatio);
    w.i32((int32_t)meshes.size());
    for (auto &m : meshes) w.i32((int32_t)settings.p_size.at(m->particleType));
    w.f64(settings.time);
    for (auto &m : meshes) {
        w.i32((int32_t)m->hierarchy.size());
        for (auto &lv : m->hierarchy) {
            w.i32((int32_t)lv.size());
            for (auto &b : lv) {
                w.i32(b.first.first);
                w.i32(b.first.second);
                w.i32(b.second.first);
                w.i32(b.second.second);
            }
        }
        for (auto &l : m->levels)
            for (auto &r : l->rectangles) {
                const size_t n = vx_plane_elems(r->n_x, r->n_p);
                for (int s = 0; s < 2; s++) w.vec(EMSolver->Download(r->desc().f[s], n));
            }
    }
    const vx_field_state &fs = EMSolver->state();
    const size_t N = fs.n_f, ne8 = (size_t)8 * fs.n_e;
    w.vec(EMSolver->Download(fs.neutral, N));
    w.vec(EMSolver->Download(fs.phi, N));
    w.vec(EMSolver->Download(fs.E, N));
    w.vec(EMSolver->Download(fs.Ex0, 1));
    for (double *a : {fs.By, fs.Bz, fs.Ey, fs.Ez, fs.Ay, fs.Az}) w.vec(EMSolver->Download(a, ne8));
    w.vec(EMSolver->Download(fs.a2, N + 1));
    if (!w.o) throw Error("checkpoint: write failed for " + path);
}

void SolverManager::LoadCheckpoint(const std::string &path) {
    DropGraphs();  // the hierarchy and plane pointers are replaced
    Reader rd;
    rd.in.open(path, std::ios::binary);
    if (!rd.in) throw Error("checkpoint: cannot open " + path);
    char magic[8];
    rd.in.read(magic, sizeof magic);
    if (!rd.in || std::memcmp(magic, kMagic, sizeof kMagic) != 0) throw Error("checkpoint: not a checkpoint file");
    if (rd.i32() != (int32_t)settings.x_size || rd.i32() != (int32_t)settings.maxDepth ||
        rd.i32() != (int32_t)settings.refinementRatio || rd.i32() != (int32_t)meshes.size())
        throw Error("checkpoint: grid settings differ from this run");
    for (auto &m : meshes)
        if (rd.i32() != (int32_t)settings.p_size.at(m->particleType)) throw Error("checkpoint: p grid differs");
    const double t = rd.f64();
    for (auto &m : meshes) {
        std::vector<level> h(rd.i32());
        for (auto &lv : h) {
            lv.resize(rd.i32());
            for (auto &b : lv) {
                b.first.first = rd.i32();
                b.first.second = rd.i32();
                b.second.first = rd.i32();
                b.second.second = rd.i32();
            }
        }
        m->Restore(h);
        m->SetFieldSolver(EMSolver.get());
        for (auto &l : m->levels)
            for (auto &r : l->rectangles) {
                const size_t n = vx_plane_elems(r->n_x, r->n_p);
                for (int s = 0; s < 2; s++) r->UploadSlot(s, rd.vec(n));
            }
    }
    const vx_field_state &fs = EMSolver->state();
    const size_t N = fs.n_f, ne8 = (size_t)8 * fs.n_e;
    auto up = [&](double *d, size_t n) {
        std::vector<double> v = rd.vec(n);
        dev->check(vx_memcpy_h2d(dev->ctx, d, v.data(), sizeof(double) * n), "restore fields");
    };
    up(fs.neutral, N);
    up(fs.phi, N);
    up(fs.E, N);
    up(fs.Ex0, 1);
    for (double *a : {fs.By, fs.Bz, fs.Ey, fs.Ez, fs.Ay, fs.Az}) up(a, ne8);
    up(fs.a2, N + 1);
    settings.time = t;
    EMSolver->RebuildMomentPlan();
}

}  // namespace vxh
