// Stand-in for the GUI's console (graphics/gui/header/console.h:11-18) used by
// the drop-in's demo, which plays the GUI: it keeps the lines it is given.
#pragma once
#include <string>
#include <vector>

class Console {
public:
    void addLine(const std::string& s) { lines.push_back(s); }
    std::vector<std::string> lines;
};
