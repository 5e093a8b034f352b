// Output writers: the reference's text files under output/ (EMSolver.cpp:
// 340-477, SolverManager.cpp:64-70 and 103-160, Mesh.cpp:877-902), same names,
// layouts and number formats.  The arrays are copied from the device on the
// calling thread; formatting and file writes run on a writer thread, so the
// time advance continues while a dump is written.  Calls are serialised: a new
// dump waits for the previous one.
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <sstream>

#include "veritas.hpp"

namespace vxh {

// Open output streams, kept across dumps as the reference keeps its ofstreams.
struct OutputFiles {
    int precision = 15;
    std::unique_ptr<std::ofstream> charge, potential, elong, etrans, bfield, asq, time;
    std::vector<std::unique_ptr<std::ofstream>> energy;

    std::ofstream &open(std::unique_ptr<std::ofstream> &s, const std::string &name) {
        if (!s) {
            std::filesystem::create_directories("output");
            s.reset(new std::ofstream("output/" + name));
            (*s) << std::scientific << std::setprecision(precision);
        }
        return *s;
    }
};

namespace {
void line(std::ofstream &o, const std::vector<double> &v, size_t n) {
    for (size_t i = 0; i < n; i++) o << v[i] << " ";
    o << std::endl;
}
}  // namespace

std::vector<double> EMFieldSolver::Download(const double *d, size_t n) const {
    std::vector<double> h(n);
    dev->check(vx_memcpy_d2h(dev->ctx, h.data(), d, sizeof(double) * n), "download");
    return h;
}

// Rectangle.cpp:284-305 (sequential order) on the device
std::vector<double> Rectangle::CalculateEnergy() {
    const size_t n = (size_t)n_p * P.rtb;
    if (!energyR_d) {
        void *d = nullptr;
        dev->check(vx_malloc(dev->ctx, sizeof(double) * n, &d), "alloc energyR");
        allocations.push_back(d);
        energyR_d = (double *)d;
    }
    dev->check(vx_energy_patch(dev->ctx, &P, energyR_d), "energy");
    std::vector<double> h(n);
    dev->check(vx_memcpy_d2h(dev->ctx, h.data(), energyR_d, sizeof(double) * n), "download energyR");
    return h;
}

// EMSolver.cpp:124-131 with Mesh.cpp:58-62 and Level.cpp:31-40, 64-78: per
// species, levels finest first, rectangles in level order
void EMFieldSolver::AssembleEnergy() {
    const size_t nsp = settings.q.size();
    energies.assign(nsp, {});
    for (size_t s = 0; s < nsp; s++) {
        energies[s].assign(settings.p_size_finest.at(s), 0.0);
        for (auto &l : meshes.at(s)->levels) {
            std::vector<double> energyL(settings.p_size_finest.at(s), 0.0);
            for (auto &r : l->rectangles) {
                std::vector<double> e = r->CalculateEnergy();
                int shift = r->p_pos, rtb = (int)r->relativeToBottom;
                for (size_t i = 0; i < e.size(); i++) energyL[shift * rtb + i] += e[i];
            }
            for (size_t i = 0; i < energyL.size(); i++) energies[s][i] += energyL[i];
        }
    }
}

// Mesh.cpp:877-902 (formatting only; the caller writes the files)
std::vector<std::pair<std::string, std::string>> Mesh::rectangleData(double tidx) const {
    std::vector<std::pair<std::string, std::string>> out;
    for (unsigned l = 0; l < levels.size(); l++) {
        if (levels[l]->rectangles.empty()) continue;
        std::stringstream name, body;
        name << "output/rectangleData/rectangleData_p" << particleType << "_l" << l << "_t" << std::scientific
             << tidx << ".txt";
        body << std::setprecision(4);
        for (auto &r : levels[l]->rectangles) {
            body << "r " << r->x_pos << " " << r->p_pos << " " << r->n_x << " " << r->n_p << "\n";
            std::vector<double> f = r->Download(r->desc().f[0], 0);
            for (int i = 0; i < r->n_x; i++)
                for (int j = 0; j < r->n_p; j++)
                    body << r->x_pos + i << " " << r->p_pos + j << " " << std::scientific << " "
                         << f[(size_t)i * r->n_p + j] << "\n";
        }
        out.emplace_back(name.str(), body.str());
    }
    return out;
}

void SolverManager::FlushOutput() {
    if (writer.joinable()) writer.join();
}

SolverManager::~SolverManager() {
    FlushOutput();
    DropGraphs();
}

// SolverManager.cpp:103-160
void SolverManager::fileOutput(double t) {
    FlushOutput();
    if (!files) {
        files = std::make_shared<OutputFiles>();
        files->precision = settings.output.precision;
    }
    const Output o = settings.output;
    const vx_field_state &fs = EMSolver->state();
    const size_t N = fs.n_f, ne = fs.n_e, nsp = settings.q.size();
    struct Snap {
        std::vector<std::vector<double>> charges, energies;
        std::vector<double> phi, E, Ey, Ez, By, Bz, a2;
        double t;
    };
    auto snap = std::make_shared<Snap>();
    snap->t = t;
    if (o.energy) {
        EMSolver->AssembleEnergy();
        snap->energies = EMSolver->energies;
    }
    if (o.charge) {
        std::vector<double> c = EMSolver->Download(EMSolver->charges_species, nsp * N);
        for (size_t s = 0; s < nsp; s++) snap->charges.emplace_back(c.begin() + s * N, c.begin() + (s + 1) * N);
    }
    if (o.potential) snap->phi = EMSolver->Download(fs.phi, N);
    if (o.EFieldLongitudinal) snap->E = EMSolver->Download(fs.E, N);
    if (o.EFieldTransverse) {
        snap->Ey = EMSolver->Download(fs.Ey, ne);  // RK slot 0 (EMSolver.cpp:395-405)
        snap->Ez = EMSolver->Download(fs.Ez, ne);
    }
    if (o.BFieldTransverse) {
        snap->By = EMSolver->Download(fs.By, ne);
        snap->Bz = EMSolver->Download(fs.Bz, ne);
    }
    if (o.AFieldSquared) snap->a2 = EMSolver->Download(fs.a2, N);  // GetASquared(i), i < N
    std::shared_ptr<OutputFiles> F = files;
    writer = std::thread([F, o, snap, N, ne, nsp]() {
        if (o.charge) {
            std::ofstream &s = F->open(F->charge, "charge.txt");
            for (size_t k = 0; k < nsp; k++) line(s, snap->charges[k], N);
        }
        if (o.energy) {
            if (F->energy.empty()) F->energy.resize(nsp);
            for (size_t k = 0; k < nsp; k++) {
                std::ofstream &s = F->open(F->energy[k], "dNdP_" + std::to_string(k) + ".txt");
                line(s, snap->energies[k], snap->energies[k].size());
            }
        }
        if (o.potential) line(F->open(F->potential, "potential.txt"), snap->phi, N);
        if (o.EFieldLongitudinal) line(F->open(F->elong, "EFieldLong.txt"), snap->E, N);
        if (o.EFieldTransverse) {
            std::ofstream &s = F->open(F->etrans, "EFieldTrans.txt");
            line(s, snap->Ey, ne);
            line(s, snap->Ez, ne);
        }
        if (o.BFieldTransverse) {
            std::ofstream &s = F->open(F->bfield, "BFieldTrans.txt");
            line(s, snap->By, ne);
            line(s, snap->Bz, ne);
        }
        if (o.AFieldSquared) line(F->open(F->asq, "ASquared.txt"), snap->a2, N);
        if (o.time) F->open(F->time, "time.txt") << snap->t << std::endl;
    });
}

// SolverManager.cpp:64-70
void SolverManager::OutputRectangles(double t) {
    if (!settings.output.rectangleData) return;
    FlushOutput();
    auto data = std::make_shared<std::vector<std::pair<std::string, std::string>>>();
    for (auto &m : meshes) {
        auto d = m->rectangleData(t);
        data->insert(data->end(), d.begin(), d.end());
    }
    writer = std::thread([data]() {
        std::filesystem::create_directories("output/rectangleData");
        for (auto &f : *data) std::ofstream(f.first) << f.second;
    });
}

}  // namespace vxh
