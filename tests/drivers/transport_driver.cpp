// A driver written like the reference's test/transport_test.cpp (same
// Settings values, test/transport_test.cpp:9-16), compiled against this
// repository's include/ directory instead of the reference's.
#include <chrono>
#include "include/mantiisApp.hpp"

int main(int argc, char* argv[]) {
    Settings settings;
    settings.simulation_type = "transport";
    settings.verbose = 10;
    settings.T = 300.0;
    settings.P = 2.0e6;
    settings.dx = 1.0e-9;
    settings.Fbody = 1.0e-8;
    settings.dump_macroscopic_every_timestep = 1.0e6;
    settings.dump_distributions_first_timestep = -1;
    auto app = std::make_unique<MantiisApp>(argc, argv, settings);
    app->init();
    app->run();
    return 0;
}
