// wilson_statistics -- the host statistics of analyze_wilson alone, for an existing
// output/wilson/correlators_by_config.csv (no GPU needed): jackknife cosh effective mass,
// flat-plateau search and "Wilson pseudoscalar mass = m +/- dm" as
// wilson/analyze_configs.cpp:36-79 prints them, the series in output/wilson/effective_mass.csv.
// Reads the analysis.* keys of ./parameters.txt.

#include "app_common.hpp"
#include "wilson_stats.hpp"

#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

int main()
{
    try {
        const app::Parameters prm;
        const app::PlateauSettings P{prm.integer("analysis.plateau_window_size"),
                                     prm.real("analysis.plateau_flatness_tolerance"),
                                     prm.integer("analysis.earliest_plateau_time")};
        std::ifstream f("output/wilson/correlators_by_config.csv");
        if (!f) throw std::runtime_error("cannot open output/wilson/correlators_by_config.csv");
        std::string line;
        std::getline(f, line);  // header: config,t0,...
        std::vector<app::Series> all;
        while (std::getline(f, line)) {
            if (line.empty()) continue;
            std::stringstream ss(line);
            std::string cell;
            std::getline(ss, cell, ',');  // config index
            app::Series c;
            while (std::getline(ss, cell, ',')) c.push_back(std::stod(cell));
            all.push_back(c);
        }
        const app::WilsonMass m = app::wilson_mass(all, P, &std::cout);
        app::write_effective_mass_csv("output/wilson/effective_mass.csv", m);
        std::cout << "Wrote:\n  output/wilson/effective_mass.csv\n";
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
}
