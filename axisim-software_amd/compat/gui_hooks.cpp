// compat/gui_hooks.cpp -- the GUI side of gui_hooks.hpp, compiled with the
// GUI's include path (its console.h and residual_plot.h declare
// Console::addLine(const std::string&), ResidualPlot::add(int, cfg map) and
// ResidualPlot::resetState(), graphics/gui/header/console.h:18,
// residual_plot.h:60-72).
#include "gui_hooks.hpp"

#include "console.h"
#include "residual_plot.h"

namespace axs::compat {

void consoleAddLine(Console* console, const std::string& line) {
    if (console) console->addLine(line);
}

void residualPlotAdd(ResidualPlot* plot, int iteration, const std::unordered_map<std::string, ConfigResidual>& cfg) {
    if (plot) plot->add(iteration, cfg);
}

void residualPlotResetState(ResidualPlot* plot) {
    if (plot) plot->resetState();
}

}  // namespace axs::compat
