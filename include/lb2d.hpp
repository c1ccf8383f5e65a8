// Drop-in counterpart of the reference's include/lb2d.hpp (lb2d.hpp:35-42):
// the umbrella header of the solver core.
#pragma once

#include "coreLB.hpp"
