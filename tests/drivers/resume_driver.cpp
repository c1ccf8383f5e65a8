// The transport driver of test/transport_test.cpp with this repository's
// resume and binary-output settings: an optional 6th argument names a dump
// of f (f.txt or f.bin) to start from; final fields are also written as
// f.bin, ux.bin, uy.bin, rho.bin.
#include <chrono>
#include "include/mantiisApp.hpp"

int main(int argc, char* argv[]) {
    Settings settings;
    settings.simulation_type = "transport";
    settings.verbose = 1000;
    settings.T = 300.0;
    settings.P = 2.0e6;
    settings.dx = 1.0e-9;
    settings.Fbody = 1.0e-8;
    settings.dump_macroscopic_every_timestep = 1.0e6;
    settings.dump_distributions_first_timestep = -1;
    settings.binary_output = true;
    if (argc > 6) settings.initial_distributions = argv[6];
    auto app = std::make_unique<MantiisApp>(argc < 6 ? argc : 6, argv, settings);
    app->init();
    app->run();
    return 0;
}
