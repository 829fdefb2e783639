// cmg_driver: runs a mechanics simulation from a reference-style XML settings file
// on the GPU path, the way CardioMechanics' main (mechanics/src/CardioMechanics/main.cpp)
// does on the CPU:   cmg_driver settings.xml [-parameter Key.Path=value ...]
// Mesh: Mesh.Tetgen.{Nodes,Elements,Bases,Surfaces,Unit}; activation: Mesh.LAT.FilePath
// (else Mesh.LAT.Offset); solver: Solver.Type = Static | NewmarkBeta.
#include <chrono>
#include <cstdio>
#include <iostream>

#include "cm_host.h"

using namespace cmhost;

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s settings.xml [-parameter key=value ...] [-steps N]\n", argv[0]);
        return 2;
    }
    try {
        ParameterMap p;
        p.ParseXML(argv[1]);
        int max_steps = -1;
        for (int i = 2; i < argc; i++) {
            std::string a = argv[i];
            if (a == "-parameter" && i + 1 < argc) {
                std::string kv = argv[++i];
                size_t eq = kv.find('=');
                if (eq == std::string::npos) throw std::runtime_error("-parameter expects key=value");
                p.Set(kv.substr(0, eq), kv.substr(eq + 1));
            } else if (a == "-steps" && i + 1 < argc) {
                max_steps = std::stoi(argv[++i]);
            }
        }
        TetgenModel m;
        const double unit = p.Get<double>("Mesh.Tetgen.Unit", 1.0);
        m.Load(p.Get<std::string>("Mesh.Tetgen.Nodes"), p.Get<std::string>("Mesh.Tetgen.Elements"),
               p.Get<std::string>("Mesh.Tetgen.Bases", ""), p.Get<std::string>("Mesh.Tetgen.Surfaces", ""), unit);
        m.CheckNodeSorting();
        const std::string lat = p.Get<std::string>("Mesh.LAT.FilePath", "null");
        if (lat != "null")
            m.LoadActivationTimes(lat);
        else
            for (auto& a : m.activation) a = p.Get<double>("Mesh.LAT.Offset", 0.0);
        CBSolverGPU solver;
        solver.Init(p, m);
        std::vector<StepRecord> rec;
        const auto t0 = std::chrono::steady_clock::now();
        solver.Run(&rec, max_steps);
        const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        for (const auto& r : rec) {
            std::printf("t=%.6f newton=%d krylov=%d |r|=%.3e wall=%.3fs", r.time, r.newton, r.krylov, r.residual,
                        r.wall);
            for (double v : r.volumes) std::printf(" volume=%.17g", v);
            for (double v : r.pressures) std::printf(" pressure=%.17g", v);
            if (r.repeats) std::printf(" repeats=%d", r.repeats);
            std::printf("\n");
        }
        std::printf("done: %zu steps in %.3f s\n", rec.size(), wall);
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "cmg_driver: %s\n", e.what());
        return 1;
    }
}
