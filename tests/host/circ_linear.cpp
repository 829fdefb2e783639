// Drives the host C++ CBCirculation on a linear cavity, V(t, p) = V0(t) + C p with
// V0(t) = Vrest - Vsqueeze sin^2(pi t / twitch) on [0, twitch], at a fixed step:
// prints "t p_est repeats" for every accepted step.  Used by tests/test_circulation.py
// to compare the C++ and Python restatements of the coupling step for step.
//   circ_linear dt stop C_m3_per_Pa Vrest Vsqueeze twitch [key=value ...]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "cm_host.h"

using namespace cmhost;

int main(int argc, char** argv) {
    if (argc < 7) return 2;
    const double dt = std::atof(argv[1]), stop = std::atof(argv[2]), C = std::atof(argv[3]);
    const double vrest = std::atof(argv[4]), vsq = std::atof(argv[5]), twitch = std::atof(argv[6]);
    ParameterMap p;
    for (int i = 7; i < argc; i++) {
        std::string kv = argv[i];
        const size_t eq = kv.find('=');
        p.Set(kv.substr(0, eq), kv.substr(eq + 1));
    }
    auto v0 = [&](double t) {
        const double s = std::sin(M_PI * t / twitch);
        return vrest - vsq * ((t >= 0.0 && t <= twitch) ? s * s : 0.0);
    };
    try {
        CBCirculation circ;
        circ.Init(p, {1});
        circ.InitVolumes({v0(0.0)}, dt);
        double time = 0.0;
        bool first = true;
        int steps = 0;
        while ((time < stop && stop - time > 1e-14) || first) {
            double t = first ? 0.0 : time + dt;
            int repeats = 0;
            while (true) {
                circ.Apply(t, dt);
                auto pr = circ.Pressures();
                const double pa = pr.count(0) ? pr[0] : 0.0;
                circ.AnalyzeResults({v0(t) + C * pa});
                const CBStatus f = circ.GetStatus();
                if (f == CBStatus::REPEAT) {
                    repeats++;
                    continue;
                }
                if (f == CBStatus::FAILED) {
                    std::printf("failed at t=%.17g\n", t);
                    return 1;
                }
                break;
            }
            std::printf("%.17g %.17g %d\n", t, circ.EstimatedPressures()[0], repeats);
            time = t;
            first = false;
            if (++steps > 100000) return 1;
        }
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "circ_linear: %s\n", e.what());
        return 1;
    }
}
