// Stand-in for the GUI's residual plot (graphics/gui/header/residual_plot.h:
// 60-72) used by the drop-in's demo: add() records the iteration and every
// enabled residual's *resVal, as the real plot's add() reads them.
#pragma once
#include <string>
#include <unordered_map>
#include <vector>

#include "solver.hpp"

class ResidualPlot {
public:
    void resetState() {
        iterations.clear();
        values.clear();
    }
    void add(int currentIteration, const std::unordered_map<std::string, ConfigResidual>& configResiduals) {
        iterations.push_back(currentIteration);
        for (const auto& [name, r] : configResiduals)
            if (r.enabled) values[name].push_back(*r.resVal);
    }
    std::vector<int> iterations;
    std::unordered_map<std::string, std::vector<double>> values;
};
