// SetupHalo.hpp -- declared in GenerateProblem.hpp (reference header name kept for drop-in includes).
#include "GenerateProblem.hpp"
