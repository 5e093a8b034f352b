// Checkpoint/restart (the reference has none): a binary file holding what
// the run's inputs must match, the hierarchy of every species, f slots 0 and
// 1 and the mixed-precision scale of every patch (whole planes, ghost rings
// included), the field state and the simulation clock.  A run restored from
// it continues bit for bit as the run that wrote it; a file of other inputs,
// or one that ends early, is rejected before the run's state changes.
#include <cstring>
#include <fstream>

#include "veritas.hpp"

namespace vxh {

namespace {
const char kMagic[8] = {'V', 'X', 'C', 'K', 'P', 'T', '0', '2'};

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

// What a checkpoint must agree on with the run that loads it: the inputs
// (Input/Particles as the C description), fMax per species (the error
// weights, Rectangle.cpp:68-74), the built-in scenario's physics when the
// hooks come from one, and the mixed-precision setting.  User hooks given as
// functions cannot be compared: a run restored with other hooks continues
// with those.
static std::vector<double> run_identity(const Settings &s, bool mixed) {
    vx_settings_desc d = s.Describe();
    std::vector<double> v = {d.minEfficiency, d.dx, d.refinementCriteria, d.cfl, d.sizeWeight, d.preLength,
                             d.postLength, (double)d.nx, (double)d.r, (double)d.Lfinest,
                             (double)d.quadratureDepth, (double)d.n_species};
    for (int i = 0; i < d.n_species; i++) {
        for (double x : {d.mass[i], d.charge[i], d.dp[i], d.pmin[i], (double)d.np[i], s.fMax.at(i)}) v.push_back(x);
    }
    v.push_back((double)s.identity.size());
    v.insert(v.end(), s.identity.begin(), s.identity.end());
    v.push_back(mixed ? 1.0 : 0.0);
    return v;
}

void SolverManager::SaveCheckpoint(const std::string &path) {
    Writer w;
    w.o.open(path, std::ios::binary);
    if (!w.o) throw Error("checkpoint: cannot open " + path);
    w.o.write(kMagic, sizeof kMagic);
    w.vec(run_identity(settings, meshes.at(0)->mixed));
    w.i32((int32_t)meshes.size());
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
        int32_t patches = 0;
        for (auto &l : m->levels) patches += (int32_t)l->rectangles.size();
        w.i32(patches);
        for (auto &l : m->levels)
            for (auto &r : l->rectangles) {
                const size_t n = vx_plane_elems(r->n_x, r->n_p);
                w.f64(r->fScale);  // the FP32 copies' scale of the mixed form
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
    Reader rd;
    rd.in.open(path, std::ios::binary);
    if (!rd.in) throw Error("checkpoint: cannot open " + path);
    char magic[8];
    rd.in.read(magic, sizeof magic);
    if (!rd.in || std::memcmp(magic, kMagic, sizeof kMagic) != 0) throw Error("checkpoint: not a checkpoint file");
    const std::vector<double> want = run_identity(settings, meshes.at(0)->mixed);
    const int32_t nid = rd.i32();
    if (nid < 0 || nid > (1 << 16)) throw Error("checkpoint: corrupt header");
    std::vector<double> got(nid);
    rd.in.read((char *)got.data(), sizeof(double) * nid);
    if (!rd.in) throw Error("checkpoint: truncated file");
    if (got != want)
        throw Error("checkpoint: written by a run of other inputs (grid, species, scenario, fMax or precision)");
    if (rd.i32() != (int32_t)meshes.size()) throw Error("checkpoint: species count differs");
    // the whole file is read into host memory before anything of this run changes
    struct Saved {
        std::vector<level> h;
        std::vector<double> scale;
        std::vector<std::vector<double>> slots;  // per patch: slot 0, slot 1
    };
    const double t = rd.f64();
    const int maxL = settings.maxDepth + 1;
    std::vector<Saved> saved(meshes.size());
    for (size_t mi = 0; mi < meshes.size(); mi++) {
        Saved &S = saved[mi];
        const int32_t nl = rd.i32();
        if (nl < 0 || nl > maxL) throw Error("checkpoint: corrupt hierarchy");
        S.h.resize(nl);
        for (auto &lv : S.h) {
            const int32_t nb = rd.i32();
            if (nb < 0 || nb > (1 << 20)) throw Error("checkpoint: corrupt hierarchy");
            lv.resize(nb);
            for (auto &b : lv) {
                b.first.first = rd.i32();
                b.first.second = rd.i32();
                b.second.first = rd.i32();
                b.second.second = rd.i32();
            }
        }
        // the patches in level order as promoteHierarchyToMesh makes them from
        // the hierarchy; their plane sizes are checked again on upload
        const int32_t patches = rd.i32();
        if (patches < 0 || patches > (1 << 20)) throw Error("checkpoint: corrupt patch count");
        for (int32_t k = 0; k < patches; k++) {
            S.scale.push_back(rd.f64());
            for (int s = 0; s < 2; s++) {
                const int32_t n = rd.i32();
                if (n < 0 || n > (1 << 28)) throw Error("checkpoint: corrupt plane");
                std::vector<double> v(n);
                rd.in.read((char *)v.data(), sizeof(double) * n);
                if (!rd.in) throw Error("checkpoint: truncated file");
                S.slots.push_back(std::move(v));
            }
        }
    }
    const vx_field_state &fs = EMSolver->state();
    const size_t N = fs.n_f, ne8 = (size_t)8 * fs.n_e;
    std::vector<std::vector<double>> fields;
    for (size_t n : {N, N, N, (size_t)1, ne8, ne8, ne8, ne8, ne8, ne8, N + 1}) fields.push_back(rd.vec(n));
    // everything is read: replace this run's state
    DropGraphs();  // the hierarchy and plane pointers are replaced
    for (size_t mi = 0; mi < meshes.size(); mi++) {
        auto &m = meshes[mi];
        Saved &S = saved[mi];
        m->Restore(S.h);
        m->SetFieldSolver(EMSolver.get());
        size_t k = 0;
        for (auto &l : m->levels)
            for (auto &r : l->rectangles) {
                if (k >= S.scale.size()) throw Error("checkpoint: corrupt file (patches differ from its hierarchy)");
                for (int s = 0; s < 2; s++) r->UploadSlot(s, S.slots.at(2 * k + s));
                r->fScale = S.scale[k];
                k++;
            }
    }
    double *dst[] = {fs.neutral, fs.phi, fs.E, fs.Ex0, fs.By, fs.Bz, fs.Ey, fs.Ez, fs.Ay, fs.Az, fs.a2};
    for (size_t i = 0; i < fields.size(); i++)
        dev->check(vx_memcpy_h2d(dev->ctx, dst[i], fields[i].data(), sizeof(double) * fields[i].size()),
                   "restore fields");
    settings.time = t;
    EMSolver->RebuildMomentPlan();
}

}  // namespace vxh
