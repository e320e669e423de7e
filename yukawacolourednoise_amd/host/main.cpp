// main.cpp -- `ycn_langevin input.toml`: drop-in for the reference executable built
// from code/src/main.cu (same input keys, same stdout, data.csv, traces.csv and
// slice.dat), running every step on the GPU through libycn.so.
#include <fstream>
#include <iostream>
#include <sstream>

#include "langevin_driver.hpp"
#include "ycn_backend.hpp"

int main(int argc, char **argv) {
    ycn::install_signal_handler();
    if (argc == 1) {
        std::cerr << "No input file provided.\nExiting.\n";
        return 1;
    }
    try {
        std::ifstream f(argv[1]);
        if (!f) {
            std::cerr << "cannot open " << argv[1] << "\n";
            return 1;
        }
        std::stringstream ss;
        ss << f.rdbuf();
        const ycn::Toml toml = ycn::Toml::parse_string(ss.str());
        const ycn::RunConfig cfg = ycn::RunConfig::from_toml(toml);
        ycn::YcnBackend backend(cfg);
        ycn::DriverIO io;
        io.out = &std::cout;
        io.dir = ".";
        io.inputEcho = ss.str();
        ycn::run_langevin(cfg, backend, io);
    } catch (const std::exception &e) {
        std::cerr << "ycn_langevin: " << e.what() << "\n";
        return 2;
    }
    return 0;
}
