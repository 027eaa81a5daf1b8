// The reference README's usage (README.md:57-99 of the reference) through
// the drop-in header.  "host" runs the calls that need no GPU and checks the
// error contract; "gpu" also starts the step and runs frames until agent 0
// arrives.
#include <cstdio>
#include <cstring>
#include <vector>

#include "gmapf/GMAPF.h"

using namespace gmapf;

static int fail(const char* what) {
    std::printf("FAIL %s\n", what);
    return 1;
}

int main(int argc, char** argv) {
    bool gpu = argc > 1 && std::strcmp(argv[1], "gpu") == 0;
    Config config;
    config.AgentsMaxCount = 2048;
    config.AgentSpeed = 1;
    config.AgentRadius = 2;
    config.PathFinderParallelAgents = 2048;
    config.PathFinderThreadsPerAgents = 4;
    config.PathFinderQueueCapacity = 16;
    config.PathFinderHeuristicK = 1;
    config.IsProfiler = true;

    GMAPF g;
    bool threw = false;
    try {
        g.AddAgent({1, 1});
    } catch (const std::exception&) {
        threw = true;
    }
    if (!threw) return fail("AddAgent before Init must throw");

    g.Init(config);
    g.FillMapStart(100, 100);
    std::vector<V2Float> wall;
    wall.emplace_back(30, 40);
    wall.emplace_back(30, 60);
    wall.emplace_back(60, 40);
    wall.emplace_back(60, 60);
    g.FillMapObstacle(wall);
    g.FillMapEnd();

    int id = g.AddAgent({1, 1});
    if (id != 0) return fail("first agent id");
    threw = false;
    try {
        g.AddAgent({35, 50});  // inside the left lobe of the wall
    } catch (const std::exception&) {
        threw = true;
    }
    if (!threw) return fail("AddAgent inside an obstacle must throw");
    g.SetAgentTargetPos(0, {99, 99});
    if (g.GetAgentState(0) != EAgentState::Search) return fail("state after SetAgentTargetPos");
    if (!gpu) {
        std::printf("host ok\n");
        return 0;
    }
    g.ManualStart();
    for (int f = 0; f < 20000; f++) {
        g.AsyncStep(1.0f / 10);
        g.WaitStepEnd();
        if (g.GetAgentState(0) == EAgentState::Idle) {
            const V2Float& p = g.GetAgentPos(0);
            float dx = p.X - 99, dy = p.Y - 99;
            if (dx * dx + dy * dy > 2.0f * 2.0f + 1e-3f) return fail("arrived too far from the target");
            std::printf("gpu ok: arrived after %d frames at (%.3f, %.3f)\n", f + 1, p.X, p.Y);
            g.ProfilerDump();
            return 0;
        }
    }
    return fail("agent did not arrive");
}
