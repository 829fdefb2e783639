// CPU self-test of the host C++ layer (no device calls): settings parsing,
// analytic functions, step-size control, Tetgen/LAT reading, material keys.
#include <cmath>
#include <cstdio>
#include <fstream>

#include "cm_host.h"

using namespace cmhost;

static int fails = 0;
#define CHECK(c)                                                  \
    do {                                                          \
        if (!(c)) {                                               \
            std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c); \
            fails++;                                              \
        }                                                         \
    } while (0)

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : ".";
    ParameterMap p;
    p.ParseXMLString(R"(<settings>Test
<!-- comment -->
<Materials>
  <Mat_Default><Density>1082</Density><Guccione><bScale>2</bScale></Guccione></Mat_Default>
  <Mat_30><Type>Guccione</Type><Guccione><C>880</C><bf>8</bf><bt>6</bt><bfs>12</bfs><K>1e6</K></Guccione>
    <TensionModel>FromFunction</TensionModel><TensionMax>60000</TensionMax>
    <FromFunction><Type>Sinus</Type><StartTime>0</StartTime><StopTime>0.3</StopTime>
      <Sinus><Omega>20.943951</Omega></Sinus></FromFunction></Mat_30>
  <Mat_7><Type>Holzapfel</Type><Holzapfel><a>330</a><b>9.24</b><af>18500</af><bf>15.97</bf><as>2560</as>
    <bs>10.45</bs><afs>420</afs><bfs>11.6</bfs><kappa>1e6</kappa></Holzapfel></Mat_7>
</Materials>
<Solver><Type>Static</Type><StopTime>1</StopTime><TimeStep>0.1</TimeStep><MinSteps>2</MinSteps>
  <Plugins><ApplyPressureFromFunction>true</ApplyPressureFromFunction></Plugins></Solver>
<Plugins><ApplyPressureFromFunction><Groups><G1><Surfaces>1,3</Surfaces><Intervals>
  <Interval_1><Type>Linear</Type><StartTime>0</StartTime><StopTime>1</StopTime><Amplitude>2000</Amplitude></Interval_1>
</Intervals></G1></Groups></ApplyPressureFromFunction></Plugins>
</settings>)");
    CHECK(p.Get<double>("Materials.Mat_30.Guccione.C") == 880.0);
    CHECK(p.Get<std::string>("Solver.Type") == "Static");
    CHECK(p.GetArray<int>("Plugins.ApplyPressureFromFunction.Groups.G1.Surfaces").size() == 2);
    CHECK(p.GetChildNodes("Materials").size() == 3);
    CHECK(!p.IsAvailable("Materials.Mat_30.Density"));

    auto mats = MaterialsFromParameters(p, {30, 7});
    CHECK(mats.size() == 2);
    CHECK(mats[0].law == CMG_LAW_GUCCIONE && mats[0].params[0] == 880.0);
    CHECK(mats[0].params[1] == 16.0 && mats[0].params[3] == 24.0);     // bScale from Mat_Default
    CHECK(mats[0].density == 1082.0 && mats[1].density == 1082.0);      // Mat_Default fallback
    CHECK(mats[0].tension_fn == CMG_FN_SINUS && mats[0].tension_max == 60000.0);
    CHECK(std::fabs(mats[0].fn[3] - 20.943951) < 1e-12);
    CHECK(mats[1].law == CMG_LAW_HOLZAPFEL && mats[1].params[8] == 1e6 && mats[1].params[9] == 100.0);

    // Land17: per-key Mat_Default fallback reads the tail-less key (no value) and throws;
    // the calcium switch has no fallback; TensionMax falls back like every material key
    {
        ParameterMap q;
        q.ParseXMLString(R"(<settings><Materials>
  <Mat_Default><TensionMax>1000</TensionMax><Land17><rateDependancy>ON</rateDependancy></Land17></Mat_Default>
  <Mat_2><Type>Guccione</Type><Guccione><C>880</C><bf>8</bf><bt>6</bt><bfs>12</bfs><K>1e6</K></Guccione>
    <TensionModel>Land17</TensionModel>
    <Land17><Tref>100</Tref><CalciumTransientType>Brixius</CalciumTransientType><TmBlocked>0.5</TmBlocked></Land17></Mat_2>
  <Mat_3><Type>Guccione</Type><Guccione><C>880</C><bf>8</bf><bt>6</bt><bfs>12</bfs><K>1e6</K></Guccione></Mat_3>
</Materials></settings>)");
        std::vector<std::vector<double>> l17;
        auto lm = MaterialsFromParameters(q, {2, 3}, &l17);
        CHECK(lm[0].tension_fn == CMG_FN_LAND17 && lm[0].tension_max == 1000.0);
        CHECK(l17.size() == 2 && l17[0].size() == CMG_LAND17_NPARAMS && l17[1].empty());
        CHECK(l17[0][16] == 100.0 && l17[0][0] == 2.0 && l17[0][24] == 1.0 && l17[0][26] == 1.0);
        CHECK(l17[0][25] == 0.0);   // rateDependancy is read from the material's own key only
        CHECK(l17[0][30] == 0.5 && l17[0][29] == 0.0);
        ParameterMap r;
        r.ParseXMLString(R"(<settings><Materials>
  <Mat_Default><Land17><Tref>90</Tref></Land17></Mat_Default>
  <Mat_2><Type>Guccione</Type><Guccione><C>880</C><bf>8</bf><bt>6</bt><bfs>12</bfs><K>1e6</K></Guccione>
    <TensionModel>Land17</TensionModel><TensionMax>1000</TensionMax></Mat_2>
</Materials></settings>)");
        bool threw = false;
        try {
            MaterialsFromParameters(r, {2}, &l17);
        } catch (const std::exception&) {
            threw = true;
        }
        CHECK(threw);
    }

    // time functions
    auto lin = CBAnalyticFunction::Create("Linear", p, "Plugins.ApplyPressureFromFunction.Groups.G1.Intervals.Interval_1");
    CHECK(std::fabs((*lin)(0.25) - 0.25) < 1e-15);
    // stateful window stop + (t - t_prev)/2 (CBAnalyticFunction.h:62): a large jump past
    // stopTime still ramps once, the next call holds the maximum
    CHECK((*lin)(1.4) == 1.4);
    CHECK((*lin)(1.5) == 1.0);
    CBApplyPressureFromFunction plug;
    plug.Init(p, {1, 3});
    plug.Apply(0.5, 0.1);
    CHECK(std::fabs(plug.Pressures()[0] - 1000.0) < 1e-9 && std::fabs(plug.Pressures()[1] - 1000.0) < 1e-9);

    // step-size control: halve on FAILED, double back after MinSteps successes
    CBTiming t;
    t.Init(p);
    t.ResetForSim();
    CHECK(t.GetTimeStep() == 0.1);
    t.SetSolverStatus(CBStatus::FAILED);
    CHECK(t.ComputeNextTimeStep() == CBStatus::SUCCESS && std::fabs(t.GetTimeStep() - 0.05) < 1e-15);
    t.SetSolverStatus(CBStatus::SUCCESS);
    t.ComputeNextTimeStep();
    t.ComputeNextTimeStep();
    CHECK(std::fabs(t.GetTimeStep() - 0.1) < 1e-15);

    // Tetgen + LAT
    {
        std::ofstream n(dir + "/t.node"), e(dir + "/t.ele"), l(dir + "/t.lat");
        n << "4 3 1 0\n1 0 0 0 7\n2 1 0 0 0\n3 0 1 0 0\n4 0 0 1 0\n";
        e << "1 4 1\n1 1 3 2 4 5\n";
        l << "id time\n1 0.02\n";
    }
    TetgenModel m;
    m.Load(dir + "/t.node", dir + "/t.ele", "", "", 1e-3);
    CHECK(m.nNodes == 4 && m.npe == 4 && m.bc[0] == 7 && m.material[0] == 5);
    CHECK(std::fabs(m.coords[3] - 1e-3) < 1e-18);
    m.LoadActivationTimes(dir + "/t.lat");
    CHECK(m.activation[0] == 0.02);
    std::printf("host selftest: %d failures\n", fails);
    return fails ? 1 : 0;
}
