// Runs MantiisApp with its Settings given on the command line, so that one
// executable serves every drop-in run of the tests (the regression inputs of
// the transport and adsorption cases, resume from a dump, binary output).
//
//   app_driver [key=value ...] -- <num_threads> <is_multiblock> <Nx> <input_folder/> <Nt>
//
// keys: type (transport | adsorption | coupled), dx, fbody, T, P, collision,
// wall, verbose, threshold, check_every, macro_every (dump period of the
// macroscopic fields), dist_first (first step of the distribution dumps),
// binary (0 | 1), init (a dump of f to start from: f.txt or f.bin).
// Everything after "--" is handed to MantiisApp as its argv.
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "include/mantiisApp.hpp"

static bool apply(Settings& s, const std::string& key, const std::string& v) {
    if (key == "type") s.simulation_type = v;
    else if (key == "dx") s.dx = std::stod(v);
    else if (key == "fbody") s.Fbody = std::stod(v);
    else if (key == "T") s.T = std::stod(v);
    else if (key == "P") s.P = std::stod(v);
    else if (key == "collision") s.collision_method = std::stoi(v);
    else if (key == "wall") s.wall_boundary = std::stoi(v);
    else if (key == "verbose") s.verbose = std::stoi(v);
    else if (key == "threshold") s.threshold = std::stod(v);
    else if (key == "check_every") s.check_every = std::stoi(v);
    else if (key == "macro_every") s.dump_macroscopic_every_timestep = std::stoi(v);
    else if (key == "dist_first") s.dump_distributions_first_timestep = std::stoi(v);
    else if (key == "binary") s.binary_output = std::stoi(v) != 0;
    else if (key == "init") s.initial_distributions = v;
    else return false;
    return true;
}

int main(int argc, char* argv[]) {
    Settings settings;
    int sep = -1;
    for (int i = 1; i < argc; i++) {
        if (std::strcmp(argv[i], "--") == 0) {
            sep = i;
            break;
        }
        const std::string kv = argv[i];
        const size_t eq = kv.find('=');
        if (eq == std::string::npos || !apply(settings, kv.substr(0, eq), kv.substr(eq + 1))) {
            std::cerr << "app_driver: bad setting '" << kv << "'\n";
            return 2;
        }
    }
    if (sep < 0) {
        std::cerr << "usage: app_driver [key=value ...] -- <num_threads> <is_multiblock> <Nx> <input_folder/> <Nt>\n";
        return 2;
    }
    std::vector<char*> app_argv{argv[0]};
    for (int i = sep + 1; i < argc; i++) app_argv.push_back(argv[i]);
    app_argv.push_back(nullptr);
    MantiisApp app((int)app_argv.size() - 1, app_argv.data(), settings);
    app.init();
    app.run();
    return 0;
}
