// Serving loop: parameter-scan runs of different sizes on one GPU at once.
// No counterpart in the reference (its driver, veritas.cpp:118-170, runs one
// simulation).  See serve.cpp.
#pragma once

#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>

#include "veritas.hpp"

namespace vxh {

struct ServeRequest {
    int id = 0;
    vx_scenario_params params{};
    int steps = 0;        // steps to run, dt = cfl x CFL bound before every step
    double cfl = 0.4;
    double arrival = 0.0; // seconds after Run() starts
    bool mixed = false;   // mixed-precision form
};

struct ServeRecord {
    int id = 0;
    double arrival = 0, admitted = 0, started = 0, finished = 0;  // seconds after Run() starts
    size_t footprint = 0;     // device bytes the run held
    double cell_updates = 0;  // cells of all patches x steps
    bool batched = false;     // its Vlasov stages ran in shared launches
    double time = 0;          // simulated time at the end
};

class Server {
public:
    // budget: device bytes that admitted runs may hold together; max_active:
    // runs stepped at once (at most 64, the batched launch's limit)
    Server(int device, size_t budget, int max_active);
    ~Server();
    void Submit(const ServeRequest &r);
    void Run();  // until every submitted request has finished
    std::vector<ServeRecord> records;  // completion order
    bool keep_state = false;           // keep each run's final f (slot 1, interior)
    std::map<int, std::vector<double>> final_f;
    size_t peak_bytes = 0;  // largest sum of the footprints held at once
    int peak_active = 0;
    long ticks = 0, batched_launches = 0;
    int builders = 3;  // threads that build admitted runs (set before Run)

private:
    struct Live {
        ServeRequest req;
        ServeRecord rec;
        std::unique_ptr<Settings> settings;
        std::unique_ptr<SolverManager> sm;
        double dt = 0.0;
        int done = 0;
        bool waiting = false;  // a step is in flight (its CFL bound pending)
        size_t bytes = 0;
        std::string error;
    };
    int device;
    size_t budget;
    int max_active;
    std::deque<ServeRequest> pending;  // by arrival
    std::unique_ptr<Device> batch;     // stream of the batched launches
    // construction of admitted runs on a worker thread
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::unique_ptr<Live>> to_build, built;
    bool stop = false;
    std::vector<std::thread> workers;
    void Build();
    void Tick(const std::vector<Live *> &ready);
};

size_t ServeEstimate(const vx_scenario_params &p);

}  // namespace vxh
