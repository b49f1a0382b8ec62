// compat/gui_hooks.hpp -- how the Solver drop-in reaches the GUI objects it
// holds pointers to (Solver::console, Solver::residualPlot).  The reference's
// solver.cpp calls Console::addLine, ResidualPlot::add and
// ResidualPlot::resetState directly (solver/source/solver.cpp:146-173,
// 1245-1266); here those three calls go through these hooks, which
// compat/gui_hooks.cpp defines by including the GUI's own console.h and
// residual_plot.h, so the drop-in library itself needs no GUI header.
#pragma once

#include <string>
#include <unordered_map>

#include "solver.hpp"

namespace axs::compat {
void consoleAddLine(Console* console, const std::string& line);
void residualPlotAdd(ResidualPlot* plot, int iteration, const std::unordered_map<std::string, ConfigResidual>& cfg);
void residualPlotResetState(ResidualPlot* plot);
}  // namespace axs::compat
